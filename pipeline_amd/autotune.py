"""Autotuner for the launch shapes of the culling kernels.

Tunes, on the device it runs on:
  cull   dpc_cull's waves_per_cu (0 = full grid, >0 = persistent streaming grid)
  batch  dpc_batch_config (objects per lane, threads per block, waves per CU,
         frusta per work item)
Every candidate is timed with HIP events over repeated launches and its masks
are compared bit for bit with a reference launch of the same work; a candidate
that differs is rejected.  Results are stored in pipeline_amd/tuning/<arch>.json,
keyed by kernel and shape bucket, and read back by best_batch_config /
best_cull_config.

    python -m pipeline_amd.autotune            # tune the standard workloads, print the table
    python -m pipeline_amd.autotune --write    # ... and store it in pipeline_amd/tuning/
"""
from __future__ import annotations

import argparse
import ctypes as C
import itertools
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import _lib, scenes
from . import device as dev

TUNING_DIR = Path(__file__).resolve().parent / "tuning"


def arch_key() -> str:
    name = dev.device_info(0)["name"]
    return name.split("(")[-1].split(":")[0].strip(") ") or "unknown"


def _bucket(x: float) -> int:
    return int(round(math.log2(max(float(x), 1.0))))


def batch_bucket(objects: int, tests: int, frusta: int) -> str:
    return f"obj2^{_bucket(objects)}-tests2^{_bucket(tests)}-frusta{int(frusta)}"


def cull_bucket(objects: int) -> str:
    return f"obj2^{_bucket(objects)}"


def load_table(arch: str | None = None) -> dict:
    arch = arch or arch_key()
    f = TUNING_DIR / f"{arch}.json"
    return json.loads(f.read_text()) if f.exists() else {}


def best_batch_config(objects: int, tests: int, frusta: int, arch: str | None = None):
    t = load_table(arch).get("batch", {})
    entry = t.get(batch_bucket(objects, tests, frusta))
    if entry is None and t:
        # nearest bucket by test count
        key = min(t, key=lambda k: abs(int(k.split("tests2^")[1].split("-")[0]) - _bucket(tests)))
        entry = t[key]
    return _lib.BatchConfig(*entry["config"]) if entry else None


def best_cull_config(objects: int, arch: str | None = None):
    t = load_table(arch).get("cull", {})
    entry = t.get(cull_bucket(objects))
    if entry is None and t:
        key = min(t, key=lambda k: abs(int(k.split("2^")[1]) - _bucket(objects)))
        entry = t[key]
    return _lib.CullConfig(entry["waves_per_cu"]) if entry else None


def _prewarm(issue, stream, seconds=0.3):
    """Untimed launches for `seconds` before the first candidate: the GPU's clocks
    ramp up over tens of milliseconds, and the candidates timed first would
    otherwise be measured slow."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(32):
            issue()
        stream.sync()


def _time(issue, stream, reps=30, warm=5):
    for _ in range(warm):
        issue()
    stream.sync()
    ev = [(dev.Event(), dev.Event()) for _ in range(reps)]
    for a, b in ev:
        a.record(stream.handle)
        issue()
        b.record(stream.handle)
    stream.sync()
    ms = sorted(a.elapsed_ms(b) for a, b in ev)
    for a, b in ev:
        a.destroy()
        b.destroy()
    return float(np.median(ms))


def tune_cull(fs: scenes.FlatScene, candidates=(0, 4, 8, 12, 16, 24, 32), copies=4, log=print):
    """Times dpc_cull on `copies` copies of the scene (so the last-level cache does not
    hold the input) for each waves_per_cu."""
    lib = _lib.load()
    stream = dev.Stream()
    sc = [dev.upload_scene(fs.lower, fs.upper, fs.matrices) for _ in range(copies)]
    masks = [dev.DeviceArray(dev.mask_words(fs.n), np.uint64) for _ in range(copies)]
    vp = dev.vp_arg(fs.vp)
    old = _lib.CullConfig()
    _lib.check(lib.dpc_cull_get_config(C.byref(old)))
    ref = None
    rows = []
    _prewarm(lambda: _lib.check(lib.dpc_cull(sc[0].box_a.ptr, sc[0].box_b.ptr, None, sc[0].n_objects, sc[0].planes.ptr,
                                             sc[0].ld, vp, masks[0].ptr, stream.handle)), stream)
    for w in candidates:
        _lib.check(lib.dpc_cull_set_config(C.byref(_lib.CullConfig(w))))
        k = [0]

        def issue():
            s = sc[k[0] % copies]
            _lib.check(lib.dpc_cull(s.box_a.ptr, s.box_b.ptr, None, s.n_objects, s.planes.ptr, s.ld, vp,
                                    masks[k[0] % copies].ptr, stream.handle))
            k[0] += 1

        ms = _time(issue, stream)
        got = masks[0].download(stream.handle)
        if ref is None:
            ref = got
        ok = bool(np.array_equal(got, ref))
        gbs = (88 * fs.n + fs.n / 8) / (ms * 1e-3) / 1e9
        rows.append({"waves_per_cu": w, "us": ms * 1e3, "gbs": gbs, "ok": ok})
        log(f"cull waves_per_cu={w:3d}: {ms * 1e3:8.2f} us  {gbs:7.0f} GB/s  {'ok' if ok else 'MISMATCH'}")
    _lib.check(lib.dpc_cull_set_config(C.byref(old)))
    for s in sc:
        s.free()
    for m in masks:
        m.free()
    stream.destroy()
    good = [r for r in rows if r["ok"]]
    best = min(good, key=lambda r: r["us"])
    return best, rows


def batch_candidates(quick=False):
    ks = (1, 2) if quick else (1, 2, 4)
    blocks = (256,) if quick else (64, 128, 256)
    waves = (8, 16, 24) if quick else (4, 8, 12, 16, 24, 32)
    chunks = (0, 4) if quick else (0, 2, 4, 8)
    for k, b, w, c in itertools.product(ks, blocks, waves, chunks):
        yield (k, b, w, c)


def tune_batch(bs: scenes.BatchedScene, candidates=None, log=print):
    lib = _lib.load()
    stream = dev.Stream()
    scene = dev.upload_scene(bs.lower, bs.upper, bs.matrices)
    d_vps = dev.DeviceArray.from_host(bs.vps.reshape(-1, 16))
    masks = dev.DeviceArray(bs.mask_words, np.uint64)
    # reference result: the per-view launches of the default plan
    plan0 = dev.batched_plan(bs)
    plan0.bind([scene])
    plan0.cull_per_view(bs.vps, masks, stream.handle)
    ref = masks.download(stream.handle)
    _prewarm(lambda: plan0.cull(d_vps, masks, stream.handle), stream)
    plan0.destroy()
    rows = []
    for cfg in (candidates or batch_candidates()):
        plan = dev.batched_plan(bs, _lib.BatchConfig(*cfg))
        plan.bind([scene])
        masks.fill_u32(0, stream.handle)

        def issue():
            plan.cull(d_vps, masks, stream.handle)

        ms = _time(issue, stream, reps=20, warm=3)
        ok = bool(np.array_equal(masks.download(stream.handle), ref))
        plan.destroy()
        rows.append({"config": list(cfg), "us": ms * 1e3, "ok": ok})
        log(f"batch {cfg}: {ms * 1e3:8.2f} us  {'ok' if ok else 'MISMATCH'}")
    scene.free()
    d_vps.free()
    masks.free()
    stream.destroy()
    good = [r for r in rows if r["ok"]]
    best = min(good, key=lambda r: r["us"])
    return best, rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--write", action="store_true", help="store the table in pipeline_amd/tuning/")
    ap.add_argument("--out", default="", help="also write the table to this path")
    ap.add_argument("--quick", action="store_true")
    args = ap.parse_args(argv)
    arch = arch_key()

    def log(m):
        print(m, flush=True)

    table = load_table(arch)
    table.setdefault("batch", {})
    table.setdefault("cull", {})
    t0 = time.time()
    for n in (1 << 20,):
        fs = scenes.flat_scene_1m(n=n)
        best, rows = tune_cull(fs, log=log)
        table["cull"][cull_bucket(n)] = {"waves_per_cu": best["waves_per_cu"], "us": best["us"], "gbs": best["gbs"],
                                         "candidates": rows}
    bs = scenes.batched_scene()
    best, rows = tune_batch(bs, list(batch_candidates(args.quick)), log=log)
    table["batch"][batch_bucket(bs.n, bs.tests, bs.vps.shape[0])] = {
        "config": best["config"], "us": best["us"], "candidates": rows}
    table["device"] = dev.device_info(0)
    table["tuned_seconds"] = time.time() - t0
    text = json.dumps(table, indent=1)
    if args.write:
        TUNING_DIR.mkdir(exist_ok=True)
        (TUNING_DIR / f"{arch}.json").write_text(text)
    if args.out:
        Path(args.out).write_text(text)
    log(f"best cull: {table['cull']}")
    log(f"best batch: {best}")


if __name__ == "__main__":
    sys.exit(main())
