#!/usr/bin/env python3
"""Benchmark of the MI355X scene-update path: object-frustum tests/sec on one
GPU, batched (one persistent launch for all frusta; its hipGraph replay is timed
beside it) versus the plain per-view launch (one launch per frustum), with 0
visibility-bit mismatches.

Workload (synthetic, seeded, FP32): 2,097,152 objects in 64 groups of
log-uniform size, 64 frusta per launch (4 shadow-cascade ortho boxes, 6 cubemap
faces, 54 perspective cameras), each frustum culling 1-8 groups.  A "step" is
one cull of every (frustum, group) pair.  Inputs are resident in HBM before the
timed region.

Also measured in the same run and reported beside the headline:
  flat      1,048,576 objects x 1 frustum (HBM-bound), GB/s over four
            scene copies so the 256 MB last-level cache does not hold the input
  tree      1,000,000-node transform tree propagated by its launch schedule (one
            tile launch), then a cull of its ~2.4M leaf objects
  serving   16 resident scenes, Poisson requests at 2k-20k req/s (1-32 frusta and a
            1-10% matrix delta each), p50/p99 latency and tests/s per rate
  cpu       the CPU oracle (a restatement of dp::culling::cpu::ManagerImpl::cull)
            on one host core over a bounded sample of the batched workload

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-extra]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0          # MI355X HBM3E, published
PEAK_FP32_TFLOPS = 157.3         # MI355X FP32 vector, published (counts packed FMA: 2 lanes x 2 flops per op)
ISSUE_CEILING_TOPS = 39.3        # 256 CU x 64 lanes x 2.4 GHz: one unfused, unpacked FP32 op per lane per clock
FLOPS_PER_OBJECT = 40            # OBB: point = lower*M (16 mul + 12 add) + 3 x (4 mul)
FLOPS_PER_TEST = 140             # 4 vec*mat (112) + 7 corner adds (28)
CMPS_PER_TEST = 48               # 8 corners x 6 clip compares
OPS_PER_OBJECT = FLOPS_PER_OBJECT
OPS_PER_TEST = FLOPS_PER_TEST + CMPS_PER_TEST  # VALU operations, compares included
BYTES_PER_OBJECT = 88            # 64 B matrix + 24 B box
GRAPH_STEPS = 10                 # steps captured in one graph for the graph-of-steps timing


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--skip-extra", action="store_true", help="only the batched and per-view lines")
    ap.add_argument("--skip-serving", action="store_true", help="every line but the serving measurement")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--pmc-probe", action="store_true",
                    help="launch only a 512 MiB device copy, the batched kernel, the flat kernel, the tree "
                         "propagation and the bounding box a few times each (for rocprofv3 --pmc passes); prints no JSON")
    ap.add_argument("--serve-only", action="store_true", help="only the serving measurement (prints its JSON)")
    ap.add_argument("--valu-probe", action="store_true",
                    help="only the issue rate of independent FP32 adds and multiplies (prints its JSON)")
    ap.add_argument("--serve-streams", type=int, default=8, help="streams of the serving measurement")
    ap.add_argument("--serve-depth", type=int, default=2, help="batches in flight per stream in the serving measurement")
    ap.add_argument("--serve-direct", action="store_true",
                    help="serving: the device reads pinned deltas directly instead of copying them")
    ap.add_argument("--serve-seconds", type=float, default=1.0, help="length of each serving-rate run")
    ap.add_argument("--tune", default="", help="batch config k,block,waves_per_cu,max_frusta")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.machine()


def log(msg):
    sys.stderr.write(f"[bench] {msg}\n")
    sys.stderr.flush()


def prewarm(lib, stream, issue, seconds=0.3):
    """Untimed work for `seconds` before any measurement: the GPU's clocks ramp
    up over tens of milliseconds, longer than a few warm-up steps take."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(32):
            issue()
        stream.sync()


def run_timed(dev, lib, stream, steps, warmup, issue):
    """Times `steps` back-to-back issues: host wall clock with the device synchronised
    on both sides, and one HIP event pair on the launch stream around all of them
    (device time per step, gaps between launches included)."""
    for _ in range(warmup):
        issue()
    stream.sync()
    e0, e1 = dev.Event(), dev.Event()
    lib.dpc_device_sync()
    t0 = time.perf_counter()
    e0.record(stream.handle)
    for _ in range(steps):
        issue()
    e1.record(stream.handle)
    stream.sync()
    lib.dpc_device_sync()
    wall = time.perf_counter() - t0
    dev_ms = e0.elapsed_ms(e1) / steps
    e0.destroy()
    e1.destroy()
    return wall, dev_ms, dev_ms


def bench_next_rows(dev, lib, stream, ts, objs, d_local, d_world, tplan, tmask):
    """The steps beside the path, on the tree workload's data: the group bounding
    box of its 2.37M objects, the result diff of its cull mask, the inverse
    transpose of its 1M world matrices, and propagation with 1 % and with all of
    the local matrices dirty.  Kernel time per call from HIP events over 50 calls; their
    parity with the oracle is what tests/test_gpu_reduce.py, test_gpu_invert.py
    and test_gpu_tree.py check."""
    from pipeline_amd import _lib

    out = {}
    n_slots = ts.n_slots
    m = objs.n_objects
    # group bounding box (ManagerBitSet::calculateBoundingBox)
    ws = dev.workspace(lib.dpc_bbox_workspace(m))
    box6 = dev.DeviceArray(8, np.float32)

    def issue_bbox():
        _lib.check(lib.dpc_bbox(objs.box_a.ptr, objs.box_b.ptr, objs.xform_addr, m, d_world.ptr, n_slots, box6.ptr,
                                ws.ptr, stream.handle))

    _, k_ms, _ = run_timed(dev, lib, stream, 50, 5, issue_bbox)
    leaves = len(np.unique(ts.obj_slot))
    b = m * (24 + 4) + 64 * leaves
    out["bbox"] = dict(objects=m, us=k_ms * 1e3, bytes=b, gbs=b / (k_ms * 1e-3) / 1e9)
    ws.free()
    box6.free()
    # result diff (ResultBitSet::updateChanged): the cull mask against a copy with 1 % of its bits flipped
    words = dev.mask_words(m)
    new_h = tmask.download(stream.handle)
    rng = np.random.default_rng(3)
    flip = np.unique(rng.integers(0, m, m // 100))
    old_h = new_h.copy()
    np.bitwise_xor.at(old_h, flip >> 6, np.left_shift(np.uint64(1), (flip & 63).astype(np.uint64)))
    d_old = dev.DeviceArray.from_host(old_h)
    d_idx = dev.DeviceArray(max(flip.size, 1), np.uint32)
    d_cnt = dev.DeviceArray(1, np.uint32)
    ws = dev.workspace(lib.dpc_result_diff_workspace(m))

    def issue_diff():
        _lib.check(lib.dpc_result_diff(d_old.ptr, tmask.ptr, m, d_idx.ptr, d_cnt.ptr, 0, ws.ptr, stream.handle))

    _, k_ms, _ = run_timed(dev, lib, stream, 50, 5, issue_diff)
    count = int(d_cnt.download(stream.handle)[0])
    b = 16 * words + 4 * flip.size
    out["result_diff"] = dict(bits=m, changed=count, changed_expected=int(flip.size), us=k_ms * 1e3, bytes=b,
                              gbs=b / (k_ms * 1e-3) / 1e9)
    ws.free()
    # the tree objects' cull alone, and the cull with the result diff against the
    # same old mask fused into it (what Manager::cull runs per frame)
    d_mask = dev.DeviceArray(words, np.uint64)
    ws = dev.workspace(lib.dpc_cull_diff_workspace(m))
    vp = dev.vp_arg(ts.vp)

    def issue_cull():
        _lib.check(lib.dpc_cull(objs.box_a.ptr, objs.box_b.ptr, objs.xform_addr, m, d_world.ptr, n_slots, vp,
                                d_mask.ptr, stream.handle))

    def issue_cull_diff():
        _lib.check(lib.dpc_cull_diff(objs.box_a.ptr, objs.box_b.ptr, objs.xform_addr, m, d_world.ptr, n_slots, vp,
                                     d_mask.ptr, d_old.ptr, d_idx.ptr, d_cnt.ptr, 0, ws.ptr, stream.handle))

    _, c_ms, _ = run_timed(dev, lib, stream, 50, 5, issue_cull)
    _, cd_ms, _ = run_timed(dev, lib, stream, 50, 5, issue_cull_diff)
    count_fused = int(d_cnt.download(stream.handle)[0])
    same = bool(np.array_equal(d_mask.download(stream.handle), new_h))
    out["cull_with_result_diff"] = dict(objects=m, cull_us=c_ms * 1e3, fused_us=cd_ms * 1e3,
                                        diff_us=(cd_ms - c_ms) * 1e3, changed=count_fused, mask_equal=same)
    for d in (d_old, d_idx, d_cnt, ws, d_mask):
        d.free()
    # inverse transpose of every world matrix (dp::math::invertTranspose)
    d_inv = dev.DeviceArray(4 * n_slots * 4, np.float32)
    d_ok = dev.DeviceArray(((n_slots + 3) // 4) * 4, np.uint8)

    def issue_inv():
        _lib.check(lib.dpc_inverse_transpose(d_world.ptr, n_slots, None, n_slots, d_inv.ptr, n_slots, d_ok.ptr,
                                             stream.handle))

    _, k_ms, _ = run_timed(dev, lib, stream, 50, 5, issue_inv)
    b = n_slots * (64 + 64 + 1)
    out["inverse_transpose"] = dict(matrices=n_slots, us=k_ms * 1e3, bytes=b, gbs=b / (k_ms * 1e-3) / 1e9)
    d_inv.free()
    d_ok.free()
    # propagation with 1 % of the local matrices dirty (the event bytes are recomputed for every slot)
    dl = np.zeros(((n_slots + 3) // 4) * 4, np.uint8)
    dl[rng.choice(np.arange(1, n_slots), n_slots // 100, replace=False)] = 1
    d_dl = dev.DeviceArray.from_host(dl)
    d_dw = dev.DeviceArray(((n_slots + 3) // 4) * 4, np.uint8)
    d_dw.fill_u32(0, stream.handle)

    def issue_dirty():
        _lib.check(lib.dpc_tree_propagate_plan(tplan, d_local.ptr, d_world.ptr, d_dl.ptr, d_dw.ptr, stream.handle))

    _, k_ms, _ = run_timed(dev, lib, stream, 50, 5, issue_dirty)
    changed = int(d_dw.download(stream.handle)[1:n_slots].astype(np.int64).sum())
    out["dirty_propagation"] = dict(nodes=n_slots - 1, dirty_local=int(dl.sum()), changed_world=changed,
                                    us=k_ms * 1e3)
    # the same with every local matrix dirty (what a first compute after a relayout sees)
    dl[1:n_slots] = 1
    d_dl.upload(dl)
    _, k_ms, _ = run_timed(dev, lib, stream, 50, 5, issue_dirty)
    out["dirty_propagation_all"] = dict(nodes=n_slots - 1, us=k_ms * 1e3)
    d_dl.free()
    d_dw.free()
    cd = out["cull_with_result_diff"]
    log("next rows: cull {:.1f} us, cull with the result diff fused {:.1f} us (+{:.1f} us, {} changed, mask "
        "{})".format(cd["cull_us"], cd["fused_us"], cd["diff_us"], cd["changed"],
                     "equal" if cd["mask_equal"] else "DIFFERS"))
    log("next rows: bbox {:.1f} us ({:.0f} GB/s), result diff {:.1f} us ({} of {} changed), inverse transpose "
        "{:.1f} us ({:.0f} GB/s), propagation with 1 % dirty {:.1f} us ({} world matrices changed)".format(
            out["bbox"]["us"], out["bbox"]["gbs"], out["result_diff"]["us"], out["result_diff"]["changed"],
            out["result_diff"]["bits"], out["inverse_transpose"]["us"], out["inverse_transpose"]["gbs"],
            out["dirty_propagation"]["us"], out["dirty_propagation"]["changed_world"])
        + ", with every local dirty {:.1f} us".format(out["dirty_propagation_all"]["us"]))
    return out


def bench_serving(args, lib, orc, rates=(2000, 5000, 10000, 20000)):
    """Serving loop: 16 resident scenes of log-uniform size in [10k, 2M] objects,
    Poisson arrivals at each rate for --serve-seconds, every request = 1-32
    random frusta + a delta of 1-10% of its scene's matrices.  Latency is the
    server's submit -> masks-in-host-memory time.  The first rate's run is
    checked against the oracle on sampled requests (host replay of every delta
    of the sampled scenes in submission order)."""
    import threading

    from pipeline_amd import _lib as _lib_mod
    from pipeline_amd import device as dev_mod
    from pipeline_amd import scenes
    from pipeline_amd.serving import PinnedArray, Server

    t0 = time.perf_counter()
    scs = scenes.serving_scenes(seed=4, n_scenes=16)
    sizes = [sc.n for sc in scs]
    log(f"serving: 16 scenes, {sum(sizes)} objects (min {min(sizes)}, max {max(sizes)}), "
        f"generated in {time.perf_counter() - t0:.1f}s")
    # delta index and matrix pools in pinned memory: the uploads are DMA transfers
    pinned = []
    for sc in scs:
        pp = PinnedArray(sc.perm.shape, np.uint32)
        pp.array[:] = sc.perm
        pm = PinnedArray(sc.pool.shape, np.float32)
        pm.array[:] = sc.pool
        pinned += [pp, pm]
        sc.perm, sc.pool = pp.array, pm.array
    # PCIe copy rates between pinned host memory and the device (what every
    # request's delta and masks cross)
    pc_bytes = 256 << 20
    ph = PinnedArray(pc_bytes // 4, np.uint32)
    pd = dev_mod.DeviceArray(pc_bytes // 4, np.uint32)
    pcie = {}
    for name, fn in (("h2d", lambda: lib.dpc_memcpy_h2d(pd.ptr, ph.array.ctypes.data, pc_bytes, None)),
                     ("d2h", lambda: lib.dpc_memcpy_d2h(ph.array.ctypes.data, pd.ptr, pc_bytes, None))):
        fn()
        lib.dpc_device_sync()
        t0 = time.perf_counter()
        for _ in range(3):
            _lib_mod.check(fn())
        lib.dpc_device_sync()
        pcie[name + "_gbs"] = 3 * pc_bytes / (time.perf_counter() - t0) / 1e9
    pd.free()
    ph.free()
    log(f"serving: pinned copies {pcie['h2d_gbs']:.1f} GB/s host->device, {pcie['d2h_gbs']:.1f} GB/s device->host")
    n_slots, slot_words = 192, 32 * ((max(sizes) + 63) // 64)
    out_pool = PinnedArray((n_slots, slot_words), np.uint64)
    streams = getattr(args, "serve_streams", 8)
    depth = getattr(args, "serve_depth", 2)
    direct = bool(getattr(args, "serve_direct", False))
    srv = Server(max_scenes=16, max_batch_requests=16, streams=streams, hbm_budget_bytes=32 << 30,
                 batches_per_stream=depth, direct_delta_reads=direct)
    ids = [srv.add_scene(sc.lower, sc.upper, sc.matrices) for sc in scs]
    base_mats = [sc.matrices.reshape(-1, 16).copy() for sc in scs]
    # warm-up (untimed): requests through every output buffer, each carrying a
    # one-matrix delta that rewrites a matrix with its current value (state
    # unchanged), so first launches, first copies on every stream and first
    # DMA into each pinned buffer are not in the measurement
    warm = scenes.serving_requests(scs, seed=99, n_requests=n_slots, rate_per_s=0.0)
    noop_idx = [PinnedArray((1,), np.uint32) for _ in scs]
    noop_mat = [PinnedArray((1, 4, 4), np.float32) for _ in scs]
    for k, sc in enumerate(scs):
        noop_idx[k].array[0] = 0
        noop_mat[k].array[0] = sc.matrices[0]
    for k0 in range(0, len(warm), 32):
        ids_w = [srv.submit(ids[r.scene], r.vps, out_pool.array[k0 + j], noop_idx[r.scene].array,
                            noop_mat[r.scene].array) for j, r in enumerate(warm[k0:k0 + 32])]
        for rid in ids_w:
            srv.wait(rid)
    pinned += noop_idx + noop_mat
    runs = []
    check = None
    switch = sys.getswitchinterval()
    sys.setswitchinterval(2e-4)  # the submit loop spins; let the collecting thread in often
    for ri, rate in enumerate(rates):
        n_req = max(int(rate * args.serve_seconds), 16)
        reqs = scenes.serving_requests(scs, seed=100 + ri, n_requests=n_req, rate_per_s=float(rate))
        # a 0.2 s pre-roll at the same rate, requests without deltas (no state
        # change), so that the measured requests find the device in steady state
        # (clocks up, queues running) rather than just out of idle
        pre = scenes.serving_requests(scs, seed=300 + ri, n_requests=max(int(rate * 0.2), 1), rate_per_s=float(rate))
        empty_idx, empty_mats = np.zeros(0, np.uint32), np.zeros((0, 4, 4), np.float32)
        pre = [scenes.ServingRequest(r.scene, r.vps, empty_idx, empty_mats, r.arrival_us) for r in pre]
        shift = pre[-1].arrival_us + 1e6 / rate
        timeline = pre + [scenes.ServingRequest(r.scene, r.vps, r.delta_idx, r.delta_mats, r.arrival_us + shift)
                          for r in reqs]
        n_pre, n_all = len(pre), len(pre) + n_req
        sample = set()
        if ri == 0:  # the last request of four scenes, checked against the oracle afterwards
            last = {}
            for i, r in enumerate(reqs):
                last[r.scene] = i
            order = sorted(last, key=lambda s: sizes[s])
            sample = {last[s] for s in order[:2] + order[-2:]}
        free = list(range(n_slots))
        cv = threading.Condition()
        ids_q = [None] * n_all
        results_all = [None] * n_all
        kept = {}

        def waiter():
            for i in range(n_all):
                with cv:
                    while ids_q[i] is None:
                        cv.wait()
                rid, slot = ids_q[i]
                results_all[i] = srv.wait(rid)
                if i - n_pre in sample:
                    kept[i - n_pre] = out_pool.array[slot, :results_all[i].mask_words].copy()
                with cv:
                    free.append(slot)
                    cv.notify_all()

        th = threading.Thread(target=waiter)
        th.start()
        blocked_us = 0.0
        start = srv.now_us() + 1000.0
        for i, r in enumerate(timeline):
            due = start + r.arrival_us
            while srv.now_us() < due:
                pass
            with cv:
                if not free:
                    tb = srv.now_us()
                    while not free:
                        cv.wait()
                    blocked_us += srv.now_us() - tb
                slot = free.pop()
            rid = srv.submit(ids[r.scene], r.vps, out_pool.array[slot], r.delta_idx, r.delta_mats)
            with cv:
                ids_q[i] = (rid, slot)
                cv.notify_all()
        t_submitted = srv.now_us()
        th.join()
        results = results_all[n_pre:]  # the measured requests
        lat = np.array([x.latency_us for x in results])
        queue = np.array([x.admit_us - x.submit_us for x in results])
        execu = np.array([x.done_us - x.admit_us for x in results])
        t_first = min(x.submit_us for x in results)
        t_last = max(x.done_us for x in results)
        tests = sum(sizes[r.scene] * r.vps.shape[0] for r in reqs)
        batches = len({x.batch for x in results})
        st_now = srv.stats()
        issue_us = st_now["issue_us"] - (runs[-1]["_issue_total"] if runs else 0.0)
        delta_bytes = sum(68 * r.delta_idx.size for r in reqs)
        mask_bytes = sum(8 * ((sizes[r.scene] + 63) // 64) * r.vps.shape[0] for r in reqs)
        span_s = (t_last - t_first) * 1e-6
        worst = np.argsort(queue)[-10:][::-1]
        worst_queue = [dict(at_ms=round((results[i].submit_us - start) / 1e3, 2), queue_us=round(float(queue[i])),
                            exec_us=round(float(execu[i])), batch=results[i].batch,
                            batch_requests=results[i].batch_requests, scene=reqs[i].scene,
                            frusta=int(reqs[i].vps.shape[0]), delta=int(reqs[i].delta_idx.size)) for i in worst]
        runs.append(dict(_issue_total=st_now["issue_us"], worst_queue=worst_queue,
                         scheduler_issue_max_us_so_far=st_now["issue_max_us"], scheduler_issue_frac=issue_us * 1e-6 / span_s,
                         scheduler_issue_part_max_us_so_far=dict(zip(("delta_copies", "delta_scatters", "descriptors", "launch", "masks"),
                                                                     [round(x, 1) for x in st_now["issue_part_max_us"]])),
                         pcie_h2d_gbs=delta_bytes / span_s / 1e9, pcie_d2h_gbs=mask_bytes / span_s / 1e9,
                         offered_rate=rate, requests=n_req, preroll_requests=n_pre,
                         achieved_submit_rate=n_req / ((t_submitted - start - shift) * 1e-6),
                         p50_us=float(np.percentile(lat, 50)), p99_us=float(np.percentile(lat, 99)),
                         p90_us=float(np.percentile(lat, 90)), max_us=float(lat.max()),
                         queue_p50_us=float(np.percentile(queue, 50)), queue_p99_us=float(np.percentile(queue, 99)),
                         exec_p50_us=float(np.percentile(execu, 50)), exec_p99_us=float(np.percentile(execu, 99)),
                         mean_us=float(lat.mean()), tests=tests,
                         tests_per_s=tests / ((t_last - t_first) * 1e-6), batches=batches,
                         mean_batch_requests=n_req / batches, submitter_blocked_us=blocked_us,
                         delta_matrices=int(sum(r.delta_idx.size for r in reqs))))
        log(f"serving @ {rate}/s: p50 {runs[-1]['p50_us']:.0f} us, p99 {runs[-1]['p99_us']:.0f} us "
            f"(queue p99 {runs[-1]['queue_p99_us']:.0f}, exec p99 {runs[-1]['exec_p99_us']:.0f}), "
            f"{runs[-1]['tests_per_s'] / 1e9:.1f} G tests/s, {batches} batches, "
            f"submit rate {runs[-1]['achieved_submit_rate']:.0f}/s, PCIe in {runs[-1]['pcie_h2d_gbs']:.1f} GB/s, "
            f"out {runs[-1]['pcie_d2h_gbs']:.1f} GB/s, scheduler issuing {runs[-1]['scheduler_issue_frac']:.0%}")
        if ri == 0:
            # oracle replay of the sampled scenes' deltas in submission order
            mism, checked_tests = 0, 0
            mats = {reqs[i].scene: base_mats[reqs[i].scene].copy() for i in sample}
            for i, r in enumerate(reqs):
                if r.scene not in mats:
                    continue
                mats[r.scene][r.delta_idx] = r.delta_mats.reshape(-1, 16)
                if i in sample:
                    sc = scs[r.scene]
                    l4, e4 = scenes.extents(sc.lower, sc.upper)
                    obbs = orc.obb_build_group(l4, e4, np.arange(sc.n), mats[r.scene])
                    words = (sc.n + 63) // 64
                    for f in range(min(4, r.vps.shape[0])):
                        want = orc.cull_obbs(obbs, r.vps[f])
                        got = kept[i][f * words:(f + 1) * words]
                        mism += int(np.unpackbits((got ^ want).view(np.uint8)).sum())
                        checked_tests += sc.n
            check = dict(sampled_requests=len(sample), checked_tests=checked_tests, mismatches=mism)
            log(f"serving oracle check: {check}")
    sys.setswitchinterval(switch)
    for r in runs:
        r.pop("_issue_total", None)
    st = srv.stats()
    srv.close()
    out_pool.free()
    for p in pinned:
        p.free()
    return dict(scenes=16, objects=sum(sizes), scene_sizes=sizes, runs=runs, oracle_check=check, pinned_copy=pcie,
                server={"streams": streams, "batches_per_stream": depth, "max_batch_requests": 16, "hbm_budget_bytes": 32 << 30,
                        "deltas": "read directly from pinned memory" if direct else
                                  "gathered on the host and copied (DMA), affine matrices as 48 bytes",
                        "delta_matrices": st["delta_matrices"], "delta_matrices_packed": st["delta_packed"],
                        "scene_bytes": st["scene_bytes"], "peak_inflight_bytes": st["peak_inflight_bytes"]})


PMC_COPY_BYTES = 512 << 20


def measured_traffic(kernel: str):
    """HBM-side bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE passes of `bench.py --pmc-probe`, each calibrated on
    a 512 MiB copy), or None when the summary is absent."""
    f = ROOT / "profiles" / "cull_batch_kernel_traffic.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
            return int(line.split(":")[1])
    return None


def pmc_probe():
    """Kernels for the counter passes: a plain device copy of PMC_COPY_BYTES
    (calibrates FETCH_SIZE / WRITE_SIZE to bytes), then the batched launch of
    the headline workload, the flat single-frustum launch, the propagation of
    the 1,000,000-node tree and the group bounding box of its 2,372,872
    objects, 5 times each."""
    from pipeline_amd import _lib, scenes
    from pipeline_amd import device as dev
    from pipeline_amd import autotune
    lib = _lib.load()
    stream = dev.Stream()
    src = dev.DeviceArray(PMC_COPY_BYTES // 4, np.uint32)
    dst = dev.DeviceArray(PMC_COPY_BYTES // 4, np.uint32)
    src.fill_u32(1, stream.handle)
    for _ in range(3):
        _lib.check(lib.dpc_memcpy_d2d(dst.ptr, src.ptr, PMC_COPY_BYTES, stream.handle))
    stream.sync()
    src.free()
    dst.free()
    bs = scenes.batched_scene()
    scene = dev.upload_scene(bs.lower, bs.upper, bs.matrices)
    plan = dev.batched_plan(bs, autotune.best_batch_config(bs.n, bs.tests, bs.vps.shape[0]))
    plan.bind([scene])
    d_vps = dev.DeviceArray.from_host(bs.vps.reshape(-1, 16))
    masks = dev.DeviceArray(bs.mask_words, np.uint64)
    for _ in range(5):
        plan.cull(d_vps, masks, stream.handle)
    stream.sync()
    scene.free()
    plan.destroy()
    fs = scenes.flat_scene_1m()
    copies = [dev.upload_scene(fs.lower, fs.upper, fs.matrices) for _ in range(4)]
    fmask = dev.DeviceArray(dev.mask_words(fs.n), np.uint64)
    vp16 = dev.vp_arg(fs.vp)
    for k in range(5):
        s = copies[k % 4]
        _lib.check(lib.dpc_cull(s.box_a.ptr, s.box_b.ptr, None, s.n_objects, s.planes.ptr, s.ld, vp16, fmask.ptr,
                                stream.handle))
    stream.sync()
    for c in copies:
        c.free()
    fmask.free()
    ts = scenes.tree_scene(1_000_000, 2)
    d_local = dev.DeviceArray.from_host(dev.matrices_to_planes(ts.local))
    w0 = np.zeros((4, ts.n_slots, 4), np.float32)
    w0[:, 0, :] = np.eye(4, dtype=np.float32)
    d_world = dev.DeviceArray.from_host(w0)
    d_dirty = dev.DeviceArray(((ts.n_slots + 3) // 4) * 4, np.uint8)
    d_dirty.fill_u32(0, stream.handle)
    tplan = C.c_void_p()
    par_h = np.ascontiguousarray(ts.parent_slot, np.uint32)
    off_h = np.ascontiguousarray(ts.level_offsets, np.uint64)
    _lib.check(lib.dpc_tree_plan_create(par_h.ctypes.data, ts.n_slots, off_h.ctypes.data, len(off_h) - 1, 0,
                                        C.byref(tplan)))
    for _ in range(5):
        _lib.check(lib.dpc_tree_propagate_plan(tplan, d_local.ptr, d_world.ptr, None, d_dirty.ptr, stream.handle))
    stream.sync()
    # group bounding box of the tree's 2,372,872 objects over the propagated world matrices
    objs = dev.upload_scene(ts.obj_lower, ts.obj_upper, np.eye(4, dtype=np.float32)[None], ts.obj_slot)
    ws = dev.workspace(lib.dpc_bbox_workspace(objs.n_objects))
    box6 = dev.DeviceArray(8, np.float32)
    for _ in range(5):
        _lib.check(lib.dpc_bbox(objs.box_a.ptr, objs.box_b.ptr, objs.xform_addr, objs.n_objects, d_world.ptr,
                                ts.n_slots, box6.ptr, ws.ptr, stream.handle))
    stream.sync()
    for d in (objs, ws, box6):
        d.free()
    lib.dpc_tree_plan_destroy(tplan)
    for d in (d_local, d_world, d_dirty):
        d.free()
    stream.destroy()
    log("pmc probe: 3 copies of 512 MiB, 5 batched launches, 5 flat launches, 5 tree propagations")


def valu_rates(lib, iters=20000):
    """T lane-op/s of independent, unfused, unpacked FP32 adds and multiplies on
    the device (dpc_probe_valu_rate): the issue rate the batched cull's
    operations are measured against, beside the 256 CU x 64 lanes x 2.4 GHz model."""
    from pipeline_amd import _lib
    out = {}
    for name, op in (("add", 0), ("mul", 1)):
        r = C.c_double()
        _lib.check(lib.dpc_probe_valu_rate(op, iters, C.byref(r), None))
        out[f"fp32_{name}_T_lane_ops_per_s"] = r.value / 1e12
    return out


def main():
    args = parse()
    if args.pmc_probe:
        pmc_probe()
        return
    if args.serve_only:
        from oracle import oracle as orc
        from pipeline_amd import _lib
        print(json.dumps(bench_serving(args, _lib.load(), orc)))
        return
    if args.valu_probe:
        from pipeline_amd import _lib
        print(json.dumps(valu_rates(_lib.load())))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (requested --gpus {args.gpus}); "
              "culling would shard by group and propagation by subtree, with no exchange", file=sys.stderr)
        sys.exit(2)

    from pipeline_amd import _lib, scenes
    from pipeline_amd import device as dev

    lib = _lib.load()
    info = dev.device_info(0)
    log(f"device {info}")
    stream = dev.Stream()
    result = {}

    # ---------------- batched multi-view (headline) ----------------
    t_gen = time.perf_counter()
    bs = scenes.batched_scene()
    log(f"batched scene: {bs.n} objects, {len(bs.group_first)} groups, {bs.vps.shape[0]} frusta, "
        f"{len(bs.pair_group)} pairs, {bs.tests} tests ({time.perf_counter() - t_gen:.1f}s)")
    scene = dev.upload_scene(bs.lower, bs.upper, bs.matrices)
    G, P = len(bs.group_first), len(bs.pair_group)
    from pipeline_amd import autotune
    cfg = None
    if args.tune:
        cfg = _lib.BatchConfig(*[int(x) for x in args.tune.split(",")])
    else:
        cfg = autotune.best_batch_config(bs.n, bs.tests, bs.vps.shape[0])
    cull_cfg = autotune.best_cull_config(1 << 20)
    if cull_cfg is not None:
        _lib.check(lib.dpc_cull_set_config(C.byref(cull_cfg)))
    tuned = {"batch": list(cfg.__getattribute__(f) for f, _ in cfg._fields_) if cfg else "library default",
             "cull_waves_per_cu": cull_cfg.waves_per_cu if cull_cfg else "library default"}
    log(f"launch shapes: {tuned}")
    plan = dev.batched_plan(bs, cfg)
    plan.bind([scene])
    d_vps = dev.DeviceArray.from_host(bs.vps.reshape(-1, 16))
    masks = dev.DeviceArray(bs.mask_words, np.uint64)
    masks_pv = dev.DeviceArray(bs.mask_words, np.uint64)

    def issue_batch():
        plan.cull(d_vps, masks, stream.handle)

    def issue_per_view():
        plan.cull_per_view(bs.vps, masks_pv, stream.handle)

    prewarm(lib, stream, issue_batch)
    # the step: one launch of the batched kernel, K of them back to back; the
    # event pair around them is the kernel's time base for the roofline
    wall_b, k_mean_b, k_med_b = run_timed(dev, lib, stream, args.steps, args.warmup, issue_batch)
    # the same step replayed from a captured hipGraph, reported beside it (on
    # this runtime a replay costs more per step than a direct launch)
    graph = dev.Graph(stream, issue_batch)
    wall_g, _, _ = run_timed(dev, lib, stream, args.steps, args.warmup, lambda: graph.launch(stream))
    # a graph holding GRAPH_STEPS steps: the runtime's cost per replay (about 5 us
    # on this stack) is paid once per GRAPH_STEPS steps
    graph_k = dev.Graph(stream, lambda: [issue_batch() for _ in range(GRAPH_STEPS)])
    reps = max(args.steps // GRAPH_STEPS, 1)
    wall_gk, _, _ = run_timed(dev, lib, stream, reps, max(args.warmup // GRAPH_STEPS, 1), lambda: graph_k.launch(stream))
    graph_k.destroy()
    wall_v, k_mean_v, k_med_v = run_timed(dev, lib, stream, args.steps, args.warmup, issue_per_view)
    batched = masks.download(stream.handle)
    per_view = masks_pv.download(stream.handle)
    mism_pv = int(np.unpackbits((batched ^ per_view).view(np.uint8)).sum())

    # sampled oracle check of the batched result (full masks of a few pairs)
    from oracle import oracle as orc
    l4, e4 = scenes.extents(bs.lower, bs.upper)
    rng = np.random.default_rng(0)
    sample = rng.choice(P, size=min(6, P), replace=False)
    mism_orc, checked = 0, 0
    for p in sample:
        f, g, off = int(bs.pair_frustum[p]), int(bs.pair_group[p]), int(bs.pair_offset[p])
        a, n = int(bs.group_first[g]), int(bs.group_count[g])
        want = orc.cull_group(l4[a:a + n], e4[a:a + n], np.arange(a, a + n), bs.matrices, bs.vps[f])
        got = batched[off:off + want.size]
        mism_orc += int(np.unpackbits((got ^ want).view(np.uint8)).sum())
        checked += n

    tests = bs.tests
    referenced = int(bs.group_count[np.unique(bs.pair_group)].sum())
    ops = OPS_PER_OBJECT * referenced + OPS_PER_TEST * tests
    hbm_bytes = BYTES_PER_OBJECT * referenced + tests / 8
    value = tests / (wall_b / args.steps)
    per_view_value = tests / (wall_v / args.steps)
    achieved_tops = ops / (k_mean_b * 1e-3) / 1e12
    probe = valu_rates(lib)
    log("VALU probe: independent FP32 adds {:.2f}, multiplies {:.2f} T lane-op/s (model {:.1f})".format(
        probe["fp32_add_T_lane_ops_per_s"], probe["fp32_mul_T_lane_ops_per_s"], ISSUE_CEILING_TOPS))
    flops = FLOPS_PER_OBJECT * referenced + FLOPS_PER_TEST * tests
    achieved_tflops = flops / (k_mean_b * 1e-3) / 1e12
    log(f"batched: {k_mean_b * 1e3:.1f} us/launch (event), {wall_b / args.steps * 1e6:.1f} us/step wall "
        f"({wall_g / args.steps * 1e6:.1f} us/step as graph replays, "
        f"{wall_gk / (reps * GRAPH_STEPS) * 1e6:.1f} us/step in graphs of {GRAPH_STEPS} steps), "
        f"{value / 1e9:.2f} G tests/s; per-view {wall_v / args.steps * 1e6:.1f} us/step, "
        f"{per_view_value / 1e9:.2f} G tests/s; mismatches vs per-view {mism_pv}, vs oracle sample {mism_orc}")
    result["batched"] = dict(value=value, ms_per_step=wall_b / args.steps * 1e3, kernel_us=k_mean_b * 1e3,
                             graph_replay_us_per_step=wall_g / args.steps * 1e6,
                             graph_of_steps=GRAPH_STEPS,
                             graph_of_steps_us_per_step=wall_gk / (reps * GRAPH_STEPS) * 1e6,
                             kernel_us_median=k_med_b * 1e3, tests=tests, objects_referenced=referenced,
                             algorithmic_ops=ops, algorithmic_bytes=hbm_bytes,
                             algorithmic_flops=flops, achieved_tflops=achieved_tflops,
                             achieved_tops=achieved_tops, achieved_gbs=hbm_bytes / (k_mean_b * 1e-3) / 1e9)
    result["per_view"] = dict(value=per_view_value, ms_per_step=wall_v / args.steps * 1e3,
                              step_event_us=k_mean_v * 1e3, launches=int(bs.vps.shape[0]))
    graph.destroy()

    # ---------------- CPU baseline (oracle, one core, bounded sample) ----------------
    t_cpu0 = time.perf_counter()
    cpu_tests, cpu_time, pairs_done, passes = 0, 0.0, 0, 0
    order = rng.permutation(P)
    while cpu_time < args.cpu_seconds:
        # One pass = the reference's per-view loop over every (frustum, group) pair:
        # GroupCPU::updateOBBs once per group (then cached while nothing is dirty),
        # then the per-object test loop of cull().
        obb_cache = {}
        for p in order:
            f, g = int(bs.pair_frustum[p]), int(bs.pair_group[p])
            a, n = int(bs.group_first[g]), int(bs.group_count[g])
            t0 = time.perf_counter()
            if g not in obb_cache:
                obb_cache[g] = orc.obb_build_group(l4[a:a + n], e4[a:a + n], np.arange(a, a + n), bs.matrices)
            orc.cull_obbs(obb_cache[g], bs.vps[f])
            cpu_time += time.perf_counter() - t0
            cpu_tests += n
            pairs_done += 1
            if cpu_time >= args.cpu_seconds:
                break
        obb_cache.clear()
        passes += 1
    cpu_value = cpu_tests / cpu_time
    log(f"cpu oracle: {cpu_tests} tests in {cpu_time:.1f}s = {cpu_value / 1e6:.2f} M tests/s "
        f"({pairs_done} pairs, {time.perf_counter() - t_cpu0:.1f}s)")
    # the same passes on every host core available (at most 16): the object
    # range of each group split into whole mask words, one block per thread
    threads = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count())
    mt_tests, mt_time, mt_pairs = 0, 0.0, 0
    while mt_time < args.cpu_seconds / 2:
        obb_cache = {}
        for p in order:
            f, g = int(bs.pair_frustum[p]), int(bs.pair_group[p])
            a, n = int(bs.group_first[g]), int(bs.group_count[g])
            t0 = time.perf_counter()
            if g not in obb_cache:
                obb_cache[g] = orc.obb_build_group_mt(l4[a:a + n], e4[a:a + n], np.arange(a, a + n), bs.matrices,
                                                      threads)
            orc.cull_obbs_mt(obb_cache[g], bs.vps[f], threads)
            mt_time += time.perf_counter() - t0
            mt_tests += n
            mt_pairs += 1
            if mt_time >= args.cpu_seconds / 2:
                break
        obb_cache.clear()
    result["cpu_all_cores"] = dict(value=mt_tests / mt_time, unit="tests/s", cores=threads, tests=mt_tests,
                                   pairs=mt_pairs, seconds=mt_time, cpu=cpu_model())
    log(f"cpu oracle on {threads} threads: {mt_tests / mt_time / 1e6:.1f} M tests/s")
    scene.free()
    for d in (d_vps, masks, masks_pv):
        d.free()
    plan.destroy()

    # ---------------- flat single frustum (HBM roofline) ----------------
    if not args.skip_extra:
        fs = scenes.flat_scene_1m()
        copies = [dev.upload_scene(fs.lower, fs.upper, fs.matrices) for _ in range(4)]
        fmask = [dev.DeviceArray(dev.mask_words(fs.n), np.uint64) for _ in range(4)]
        vp16 = dev.vp_arg(fs.vp)
        it = [0]

        def issue_flat():
            k = it[0] % 4
            it[0] += 1
            s = copies[k]
            _lib.check(lib.dpc_cull(s.box_a.ptr, s.box_b.ptr, None, s.n_objects, s.planes.ptr, s.ld, vp16,
                                    fmask[k].ptr, stream.handle))

        prewarm(lib, stream, issue_flat)  # (the CPU baseline before this left the GPU idle)
        wall_f, k_mean_f, k_med_f = run_timed(dev, lib, stream, args.steps, args.warmup, issue_flat)
        l4f, e4f = scenes.extents(fs.lower, fs.upper)
        want = orc.cull_group(l4f, e4f, np.arange(fs.n), fs.matrices, fs.vp)
        got = fmask[0].download(stream.handle)
        mism_f = int(np.unpackbits((got ^ want).view(np.uint8)).sum())
        flat_bytes = BYTES_PER_OBJECT * fs.n + fs.n / 8
        result["flat"] = dict(objects=fs.n, kernel_us=k_mean_f * 1e3, kernel_us_median=k_med_f * 1e3,
                              tests_per_s=fs.n / (wall_f / args.steps), achieved_gbs=flat_bytes / (k_mean_f * 1e-3) / 1e9,
                              frac_hbm=flat_bytes / (k_mean_f * 1e-3) / 1e9 / PEAK_HBM_GBS, mismatches=mism_f)
        log(f"flat: {k_mean_f * 1e3:.2f} us/launch, {result['flat']['achieved_gbs']:.0f} GB/s, mismatches {mism_f}")
        for s in copies:
            s.free()
        for m in fmask:
            m.free()

        # ---------------- tree propagate + cull ----------------
        ts = scenes.tree_scene(1_000_000, 2)
        n_slots = ts.n_slots
        d_local = dev.DeviceArray.from_host(dev.matrices_to_planes(ts.local))
        w0 = np.zeros((4, n_slots, 4), np.float32)
        w0[:, 0, :] = np.eye(4, dtype=np.float32)
        d_world = dev.DeviceArray.from_host(w0)
        d_dirty = dev.DeviceArray(((n_slots + 3) // 4) * 4, np.uint8)
        d_dirty.fill_u32(0)
        objs = dev.upload_scene(ts.obj_lower, ts.obj_upper, np.eye(4, dtype=np.float32)[None], ts.obj_slot)
        tmask = dev.DeviceArray(dev.mask_words(objs.n_objects), np.uint64)
        tvp = dev.vp_arg(ts.vp)

        # the launch schedule of this layout (tile runs for the deep levels)
        tplan = C.c_void_p()
        par_h = np.ascontiguousarray(ts.parent_slot, np.uint32)
        off_h = np.ascontiguousarray(ts.level_offsets, np.uint64)
        _lib.check(lib.dpc_tree_plan_create(par_h.ctypes.data, n_slots, off_h.ctypes.data, len(off_h) - 1, 0,
                                            C.byref(tplan)))
        t_launches, t_runs, t_tiles = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _lib.check(lib.dpc_tree_plan_info(tplan, C.byref(t_launches), C.byref(t_runs), C.byref(t_tiles)))
        log(f"tree plan: {t_launches.value} launches ({t_runs.value} tile runs, {t_tiles.value} tiles) "
            f"for {len(off_h) - 1} levels")

        def issue_tree():
            _lib.check(lib.dpc_tree_propagate_plan(tplan, d_local.ptr, d_world.ptr, None, d_dirty.ptr, stream.handle))
            _lib.check(lib.dpc_cull(objs.box_a.ptr, objs.box_b.ptr, objs.xform.ptr, objs.n_objects, d_world.ptr,
                                    n_slots, tvp, tmask.ptr, stream.handle))

        t_steps = max(args.steps // 4, 5)
        prewarm(lib, stream, issue_tree)
        wall_t, k_mean_t, _ = run_timed(dev, lib, stream, t_steps, args.warmup, issue_tree)
        tgraph = dev.Graph(stream, issue_tree)
        wall_tg, _, _ = run_timed(dev, lib, stream, t_steps, args.warmup, lambda: tgraph.launch(stream))
        tgraph.destroy()
        tgraph_k = dev.Graph(stream, lambda: [issue_tree() for _ in range(GRAPH_STEPS)])
        t_reps = max(t_steps // GRAPH_STEPS, 1)
        wall_tgk, _, _ = run_timed(dev, lib, stream, t_reps, 1, lambda: tgraph_k.launch(stream))
        tgraph_k.destroy()
        # parity of this run's results: world matrices vs the FP32 oracle (ulp) and
        # an FP64 evaluation of the same recurrence; the cull mask vs the oracle
        got_w = np.transpose(d_world.download(stream.handle), (1, 0, 2))
        ref_off, ref_entries = ts.reference_levels()
        loc = np.ascontiguousarray(ts.local_by_index().reshape(-1, 16))
        w32 = np.zeros_like(loc)
        w32[0] = np.eye(4, dtype=np.float32).reshape(16)
        orc.tree_compute(ref_off, ref_entries, loc, w32, np.ones(n_slots, np.uint8), np.zeros(n_slots, np.uint8))
        want_w = w32.reshape(-1, 4, 4)[ts.index_of_slot]
        ia = got_w[1:].view(np.int32).astype(np.int64)
        ib = want_w[1:].view(np.int32).astype(np.int64)
        ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
        ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
        max_ulp = int(np.abs(ia - ib).max())
        w64 = orc.tree_compute_f64(ref_off, ref_entries, ts.local_by_index().astype(np.float64))[ts.index_of_slot]
        scale = np.abs(w64[1:]).max(axis=(1, 2), keepdims=True)
        max_rel = float((np.abs(got_w[1:].astype(np.float64) - w64[1:]) / scale).max())
        l4t, e4t = scenes.extents(ts.obj_lower, ts.obj_upper)
        want_t = orc.cull_group(l4t, e4t, ts.obj_slot, want_w, ts.vp)
        got_t = tmask.download(stream.handle)
        mism_t = int(np.unpackbits((got_t ^ want_t).view(np.uint8)).sum())
        log(f"tree parity: max ulp vs FP32 oracle {max_ulp}, max relative error vs FP64 {max_rel:.2e}, "
            f"cull mismatches {mism_t}")
        n_nodes = n_slots - 1
        m_obj = objs.n_objects
        prop_bytes = n_nodes * (64 + 64 + 4) + 64 * (n_slots - int(np.diff(ts.level_offsets)[-1]))
        cull_bytes = m_obj * (24 + 4) + 64 * len(np.unique(ts.obj_slot)) + m_obj / 8
        result["tree"] = dict(nodes=n_nodes, levels=len(ts.level_offsets) - 1, objects=m_obj,
                              us_per_step=wall_t / t_steps * 1e6, step_event_us=k_mean_t * 1e3,
                              graph_replay_us_per_step=wall_tg / t_steps * 1e6,
                              graph_of_steps_us_per_step=wall_tgk / (t_reps * GRAPH_STEPS) * 1e6,
                              achieved_gbs=(prop_bytes + cull_bytes) / (k_mean_t * 1e-3) / 1e9,
                              tests_per_s=m_obj / (wall_t / t_steps),
                              world_max_ulp_vs_oracle=max_ulp, world_max_rel_err_vs_fp64=max_rel,
                              cull_mismatches=mism_t,
                              propagation_launches=t_launches.value, tile_runs=t_runs.value, tiles=t_tiles.value)
        log(f"tree: {result['tree']['us_per_step']:.1f} us/step (propagate + cull; "
            f"{result['tree']['graph_replay_us_per_step']:.1f} us as graph replays, "
            f"{result['tree']['graph_of_steps_us_per_step']:.1f} us in graphs of {GRAPH_STEPS} steps)")
        result["next_rows"] = bench_next_rows(dev, lib, stream, ts, objs, d_local, d_world, tplan, tmask)
        for d in (d_local, d_world, d_dirty, tmask):
            d.free()
        objs.free()
        lib.dpc_tree_plan_destroy(tplan)

        # ---------------- serving loop ----------------
        if not args.skip_serving:
            result["serving"] = bench_serving(args, lib, orc)

    stream.destroy()
    line = {
        "metric": "object-frustum tests/sec (1 GPU), batched vs per-view, 0 bit mismatches",
        "value": value,
        "unit": "tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_b / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generators in pipeline_amd/scenes.py)",
        "config": {"workload": "batched multi-view cull: 2,097,152 objects in 64 log-uniform groups x 64 frusta "
                               "(4 cascades, 6 cube faces, 54 cameras), 1-8 groups per frustum",
                   "objects": bs.n, "groups": G, "frusta": int(bs.vps.shape[0]), "pairs": P, "tests": tests,
                   "parallelism": "single GPU"},
        "mismatches": {"batched_vs_per_view_bits": mism_pv, "batched_vs_oracle_sampled_bits": mism_orc,
                       "oracle_sampled_tests": checked},
        "speedup_vs_per_view": value / per_view_value,
        # The dominant kernel against MI355X's published FP32 vector peak, counting the
        # algorithmic flops (40 per object + 140 per test; compares are not flops).
        # Bit parity with the reference forbids fused multiply-add (it rounds add and
        # mul separately) and this pool forbids packed FP32, so one lane does one
        # unfused op per clock: the issue ceiling below is a quarter of the published
        # peak, and it counts the 48 compares per test as the VALU operations they are
        # (compares issue faster than adds and muls, so the kernel can reach 1.0 of it).
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": measured_traffic("cull_batch_kernel"),
                     "traffic_unit": "bytes per launch beyond L2 (FETCH_SIZE + WRITE_SIZE, copy-calibrated; "
                                     "profiles/cull_batch_kernel_traffic.md)",
                     "kernel": "cull_batch_kernel", "kernel_us": k_mean_b * 1e3,
                     "basis": f"{FLOPS_PER_OBJECT} flops/object + {FLOPS_PER_TEST} flops/test (FP32 add/mul) "
                              "over the event-timed kernel; peak = published FP32 vector peak",
                     "issue_ceiling": {"achieved": achieved_tops, "peak": ISSUE_CEILING_TOPS, "unit": "T VALU op/s",
                                       "frac": achieved_tops / ISSUE_CEILING_TOPS,
                                       "measured_fp32_add_mul": probe,
                                       "basis": f"{OPS_PER_OBJECT} ops/object + {OPS_PER_TEST} ops/test (add, mul "
                                                "and compare); 256 CU x 64 lanes x 2.4 GHz, no FMA, no packed FP32. "
                                                "Independent FP32 adds or muls alone issue at the measured rates "
                                                "below; compares issue faster, so about 1.0 is the issue limit of this "
                                                "instruction mix (profiles/cull_batch_kernel_valu.md)"},
                     "hbm": {"achieved": hbm_bytes / (k_mean_b * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": hbm_bytes / (k_mean_b * 1e-3) / 1e9 / PEAK_HBM_GBS,
                             "basis": f"{BYTES_PER_OBJECT} B per referenced object + 1 bit per test"}},
        "cpu_baseline": {"value": cpu_value, "unit": "tests/s", "cores": 1, "kind": "port",
                         "sample": f"{cpu_tests} tests: {pairs_done} (frustum, group) pairs over {passes} pass(es) of this "
                                   "workload, OBBs built once per group and pass, then the per-object test loop; "
                                   "gcc -O2 -ffp-contract=off for the x86-64 baseline ISA (no AVX-512, no FMA)"},
        "launch_shapes": tuned,
        "device": info,
        "host": {"cpu": cpu_model(), "nproc": os.cpu_count(),
                 "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None},
        "extra": result,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
