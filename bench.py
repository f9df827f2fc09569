"""Benchmark of the GPU-driven frame on one MI355X.

One step = one whole frame of the large workload through the C ABI
(lh_frame): reset, instance cull + 64-bit-atomic prefix sum, meshlet cull,
visibility-buffer raster at 1920x1080, depth bounds, cascade fitting, shadow
instance/meshlet cull and raster of 4 cascades at 2048^2, shading resolve with
fused tonemap.  Inputs are resident in HBM before timing starts.

Prints one JSON line (rank 0, N = 1).  `value` is ms per frame (lower is
better); Gtri/s uses the visible-candidate triangles (triangles of the
instances surviving the main-view instance cull).  `roofline` is for the
kernel with the largest device time per frame, timed live with HIP events
recorded on the frame's stream; `cpu_baseline` times the scalar oracle on
host cores over a bounded sample of the same frames.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "ms/frame & Gtri/s (cull+raster+shade) 10M tris 1080p; batched views/s vs plain"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=["large", "small", "stress"], default="large")
    ap.add_argument("--cpu-frames", type=int, default=10, help="oracle frames timed for the CPU baseline (0 = skip)")
    ap.add_argument("--batch-steps", type=int, default=10, help="batched-views timing iterations (0 = skip)")
    ap.add_argument("--diff-steps", type=int, default=20, help="differentiable-resolve timing iterations (0 = skip)")
    ap.add_argument("--profile-only", action="store_true", help="run a few frames for an external profiler")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run 256 MiB device-to-device copies (HBM-counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--no-atomic-probe", action="store_true", help="skip the atomicMin ceiling measurement")
    ap.add_argument("--quantized-steps", type=int, default=100,
                    help="frames timed on the same workload in the reference's quantized vertex formats (0 = skip)")
    ap.add_argument("--bc1-steps", type=int, default=100,
                    help="frames timed on the same workload with BC1-compressed textures (0 = skip)")
    ap.add_argument("--hiz", action="store_true",
                    help="time the frame with two-phase occlusion culling (output is identical either way; the "
                         "other setting is timed too)")
    ap.add_argument("--diff-only", action="store_true",
                    help="only the differentiable-resolve measurement (JSON), e.g. under a profiler")
    ap.add_argument("--batch-profile", action="store_true",
                    help="per-kernel device times of the 64-camera batch, per-view loop vs batched (JSON)")
    return ap.parse_args()


COPY_PROBE_BYTES = 256 << 20


def probe_lib():
    import ctypes as C

    path = ROOT / "lighthugger_amd" / "liblh_probe.so"
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build first (__graft_entry__.build())")
    L = C.CDLL(str(path))
    for fn in (L.lhp_atomic_min_u64, L.lhp_atomic_min_u32):
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    L.lhp_copy.restype = C.c_int
    L.lhp_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
    return L


def copy_probe(steps: int) -> None:
    """lhp_copy of COPY_PROBE_BYTES: the known-traffic kernel (k_copy) that
    scales rocprofv3's FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch

    L = probe_lib()
    a = torch.ones(COPY_PROBE_BYTES, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(max(1, steps)):
        if L.lhp_copy(a.data_ptr(), b.data_ptr(), COPY_PROBE_BYTES, stream):
            raise RuntimeError("lhp_copy failed")
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    print(f"# copy probe: {steps} copies of {COPY_PROBE_BYTES} bytes", file=sys.stderr)


def atomic_ceiling(width: int, height: int, bits: int) -> dict:
    """Device-scope atomicMin throughput (G ops/s) on a width x height buffer of
    `bits`-bit words: lanes of a wave on one 8x8 tile (the wave-cooperative and
    big-triangle shape) and on scattered pixels (one triangle per lane)."""
    import torch

    L = probe_lib()
    fn = L.lhp_atomic_min_u64 if bits == 64 else L.lhp_atomic_min_u32
    dt = torch.int64 if bits == 64 else torch.int32
    buf = torch.full((height * width,), -1 if bits == 32 else (1 << 62), dtype=dt, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    waves, ops = 1 << 16, 64
    out = {}
    for name, pattern in (("tile8x8", 0), ("scattered", 1)):
        if fn(buf.data_ptr(), width, height, waves, ops, pattern, stream):
            raise RuntimeError("atomic probe launch failed")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(5):
            fn(buf.data_ptr(), width, height, waves, ops, pattern, stream)
        e1.record()
        torch.cuda.synchronize()
        out[name] = round(5 * waves * 64 * ops / (e0.elapsed_time(e1) * 1e-3) / 1e9, 2)
    return out


def meshlet_sizes(scene):
    """(index_count, triangle_count) for every (mesh_info, meshlet), with per-mesh-info offsets."""
    offs, nv, nt = [0], [], []
    for i in range(len(scene.mesh_infos)):
        ml = scene.meshlets(i)
        nv.append(ml["index_count"].astype(np.int64))
        nt.append(ml["triangle_count"].astype(np.int64))
        offs.append(offs[-1] + len(ml))
    return np.array(offs[:-1]), np.concatenate(nv), np.concatenate(nt)


def raster_algorithmic_bytes(scene, draws, covered_pixels, sizes):
    """Compulsory HBM bytes of the per-meshlet raster of one pass: per surviving
    meshlet the draw record (8), meshlet header (56), stable-id base (4), remap
    (4 nv), FP32 positions (12 nv) and micro indices (3 nt, 4-byte padded); per
    distinct instance its Instance (112) and MeshInfo (104) records; one 8-byte
    key per covered pixel."""
    offs, nv_all, nt_all = sizes
    inst = draws["instance_index"].astype(np.int64)
    mi = scene.instances["mesh_info_address"].astype(np.int64)[inst]
    g = offs[mi] + draws["meshlet_index"].astype(np.int64)
    nv, nt = nv_all[g], nt_all[g]
    per_meshlet = 8 + 56 + 4 + 16 * nv + ((3 * nt + 3) & ~3)
    return int(per_meshlet.sum() + 216 * len(np.unique(inst)) + 8 * covered_pixels)


def tile_algorithmic_bytes(scene, view, sizes, covered_pixels):
    """Compulsory HBM bytes of one launch of the tile raster (k_tile_raster):
    per (meshlet, tile) entry the meshlet's snapped vertices from the bin cache
    (12 per vertex), its triangles of the entry (4 each), the header words and
    the entry itself (12); one 8-byte key per covered pixel written."""
    from lighthugger_amd import abi

    cfg = view.cfg
    L = abi.scratch_layout(cfg.width, cfg.height, view.dcap, cfg.big_capacity, cfg.clip_capacity, view.bcap)
    base = view.offsets["scratch"]

    def words(off, count):
        return view.mem.read(view.handle, base + off, 4 * count).view(np.uint32)

    n_items = int(view.counters()[abi.CNT_TILE_ITEMS])
    if n_items == 0:
        return 8 * covered_pixels
    items = words(L["items"], 3 * n_items).reshape(-1, 3)
    n_entries = int((items[:, 2] & 0x7FFFFFFF).max())
    ent = words(L["entries"], n_entries)
    g = (ent & ((1 << 25) - 1)).astype(np.int64)
    part = (ent >> 25).astype(np.int64)
    hdr = words(L["hdr"], 4 * (int(g.max()) + 1)).reshape(-1, 4)
    field = hdr[g, 1].astype(np.int64)
    ntri, lp = field & 255, (field >> 8) & 7
    count = (((part + 1) * ntri) >> lp) - ((part * ntri) >> lp)
    offs, nv_all, _ = sizes
    draws = view.draws(0)[g]
    mi = scene.instances["mesh_info_address"].astype(np.int64)[draws["instance_index"].astype(np.int64)]
    nv = nv_all[offs[mi] + draws["meshlet_index"].astype(np.int64)]
    return int((12 * nv + 4 * count + 12).sum() + 8 * covered_pixels)


def resolve_algorithmic_bytes(scene, view, covered_pixels):
    """Compulsory HBM bytes of one resolve launch, a lower bound: per pixel its
    key (8), FP32 radiance (16) and BGRA8 (4); per distinct vertex the covered
    pixels' triangles use, its position, normal, uv (32) and remap index (4).
    Texel and shadow-map reads are not counted."""
    from lighthugger_amd import abi

    keys = view.visbuffer().ravel()
    keys = keys[keys != np.uint64(abi.LH_VISBUFFER_CLEAR)]
    mid = ((keys >> np.uint64(abi.LH_KEY_TRI_BITS)) & np.uint64((1 << 25) - 1)).astype(np.int64)
    tri = (keys & np.uint64((1 << abi.LH_KEY_TRI_BITS) - 1)).astype(np.int64)
    ref = scene.meshlet_table[mid]
    mi = scene.instances["mesh_info_address"].astype(np.int64)[ref["instance_index"].astype(np.int64)]
    n_vertices = 0
    for m in np.unique(mi):
        sel = mi == m
        ml = scene.meshlets(int(m))[ref["meshlet_index"][sel].astype(np.int64)]
        nt = int((ml["triangle_offset"].astype(np.int64) + 3 * ml["triangle_count"].astype(np.int64)).max())
        micro = scene.micro(int(m), nt).astype(np.int64)
        corners = micro[ml["triangle_offset"].astype(np.int64)[:, None] + 3 * tri[sel][:, None] + np.arange(3)]
        idx = ml["index_offset"].astype(np.int64)[:, None] + corners
        n_vertices += len(np.unique(idx))
    pixels = view.cfg.width * view.cfg.height
    return int(pixels * 28 + 36 * n_vertices)


def batch_bench(scene, steps: int, n_cameras: int = 64):
    """Batched views: n_cameras seeded square cameras (side log-uniform in
    [256, 2048], multiples of 8), each with its own 4 cascades of the same side
    (5 views per camera).  Times the per-view loop and the batched launch, each
    with direct launches and as a replayed hipGraph, and checks on the device
    that every view's outputs are bit-identical between loop and batch."""
    import torch

    from lighthugger_amd.lib import BATCH_BATCHED, BATCH_LOOP, BatchRenderer
    from lighthugger_amd.scene import batch_cameras

    from lighthugger_amd import abi

    cams = batch_cameras(4, n_cameras, 256, 2048)
    br = BatchRenderer(scene, cams, big_capacity=1 << 20)
    n_views = n_cameras * 5

    def overflow() -> int:
        return int(np.bitwise_or.reduce([int(v.counters()[abi.CNT_OVERFLOW]) for v in br.views]))

    def outputs():
        out = []
        for v in br.views:
            t = v.handle[0]
            for name in ("visbuffer", "shadow", "radiance", "ldr"):
                o, n = v.offsets[name], v.sizes[name]
                out.append(t[o: o + n])
        return out

    rates = {}
    for name, mode, graph in (("loop", BATCH_LOOP, False), ("loop_graph", BATCH_LOOP, True),
                              ("batched", BATCH_BATCHED, False), ("batched_graph", BATCH_BATCHED, True)):
        br.run(mode, graph)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            br.run(mode, graph)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        rates[name] = {"ms_per_batch": round(dt * 1000, 3), "views_per_s": round(n_views / dt, 1)}
        ovf = overflow()
        if ovf:
            raise RuntimeError(f"batched views: capacity overflow {ovf:#x} in mode {name}")
        if name == "loop":
            ref = [o.clone() for o in outputs()]
        elif name in ("batched", "batched_graph"):
            same = all(torch.equal(a, b) for a, b in zip(ref, outputs()))
            rates[name]["bit_identical_to_loop"] = bool(same)
    br.close()
    return {
        "cameras": n_cameras, "views": n_views, "sides": [int(c.width) for c in cams],
        "pixels_per_batch": int(sum(c.width * c.height * 5 for c in cams)),
        "modes": rates,
        "speedup_vs_loop": round(rates["batched_graph"]["views_per_s"] / rates["loop"]["views_per_s"], 3),
        "speedup_vs_loop_graph": round(rates["batched_graph"]["views_per_s"] / rates["loop_graph"]["views_per_s"], 3),
    }


def diff_bench(steps: int, tri_scale_pct: int = 18, side: int = 1024) -> dict:
    """Differentiable resolve: the large scene at tri_scale_pct% of its
    triangles (about 1M visible-candidate triangles), one square view, 4
    cascades of the same side.  Times the parameterised forward resolve
    (lh_resolve_ex) and the adjoint (lh_resolve_backward, gradients of every
    parameter) with HIP events on the stream they run on."""
    import torch

    from lighthugger_amd import abi
    from lighthugger_amd.diff import DifferentiableResolve
    from lighthugger_amd.frame import ViewConfig
    from lighthugger_amd.lib import Renderer
    from lighthugger_amd.scene import HostScene, make_camera, preset_camera

    scene = HostScene.preset(abi.PRESET_LARGE, tri_scale_pct=tri_scale_pct)
    base = preset_camera(abi.PRESET_LARGE)
    r = Renderer(scene, ViewConfig(side, side, side, big_capacity=1 << 20))
    r.set_camera(make_camera(base.position.tolist(), base.target.tolist(), side, side))
    r.frame()
    torch.cuda.synchronize()
    _, _, vis0 = r.view.prefix(0)
    candidates = int(scene.instance_triangles()[vis0["index"].astype(np.int64)].sum())
    covered = int((r.view.visbuffer() != np.uint64(abi.LH_VISBUFFER_CLEAR)).sum())
    d = DifferentiableResolve(r)
    p = d.parameters()
    g = torch.randn(side, side, 4, device="cuda")
    out = d.raw_backward(p, g)

    def timed(fn) -> float:
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    def backward():
        for t in out.values():
            t.zero_()
        d.raw_backward(p, g, out, recorded=True)  # fixed quantities recorded by the forward

    fwd = timed(lambda: d.forward_raw(p))
    bwd = timed(backward)
    zero = timed(lambda: [t.zero_() for t in out.values()])
    unsupported = d.unsupported()
    return {"workload": f"large scene at {tri_scale_pct}% triangles, {side}x{side} + 4x{side}^2, visibility fixed",
            "scene_triangles": int(scene.total_triangles), "visible_candidate_triangles": candidates,
            "covered_pixels": covered, "forward_ms": round(fwd, 4), "backward_ms": round(bwd - zero, 4),
            "gradient_zeroing_ms": round(zero, 4), "backward_over_forward": round((bwd - zero) / fwd, 3),
            "unsupported_pixels": unsupported}


def variant_bench(preset: int, cfg, cam, steps: int, warmup: int, **scene_kw) -> dict:
    """The same frame with the scene stored differently (scene_kw of
    HostScene.preset: the reference's quantized vertex formats, or BC1
    textures): wall-clock ms per frame over `steps` frames, like the headline
    number."""
    import torch

    from lighthugger_amd import abi
    from lighthugger_amd.lib import Renderer
    from lighthugger_amd.scene import HostScene

    scene = HostScene.preset(preset, **scene_kw)
    r = Renderer(scene, cfg)
    r.set_camera(cam)
    for _ in range(max(1, warmup)):
        r.frame()
    torch.cuda.synchronize()
    if r.status()[abi.CNT_OVERFLOW]:
        raise RuntimeError(f"capacity overflow in the frame with {scene_kw}")
    t0 = time.perf_counter()
    for _ in range(steps):
        r.frame()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    _, _, vis0 = r.view.prefix(0)
    candidates = int(scene.instance_triangles()[vis0["index"].astype(np.int64)].sum())
    return {"ms_per_frame": round(ms, 4), "gtri_per_s": round(candidates / (ms * 1e-3) / 1e9, 3),
            "visible_candidate_triangles": candidates, "geometry_bytes": int(scene.geometry.nbytes),
            "texel_bytes": int(scene.texels.nbytes)}


def batch_profile(scene, n_cameras: int = 64, repeats: int = 3) -> dict:
    """Per-kernel device time (ms, summed over launches, mean of `repeats`
    frames) of the batch in both modes (lh_batch_profiled)."""
    import torch

    from lighthugger_amd import abi
    from lighthugger_amd.lib import BATCH_BATCHED, BATCH_LOOP, BatchRenderer
    from lighthugger_amd.scene import batch_cameras

    br = BatchRenderer(scene, batch_cameras(4, n_cameras, 256, 2048), big_capacity=1 << 20)
    out = {"cameras": n_cameras, "views": 5 * n_cameras}
    for name, mode in (("loop", BATCH_LOOP), ("batched", BATCH_BATCHED)):
        br.run(mode)
        torch.cuda.synchronize()
        acc: dict[str, float] = {}
        for _ in range(repeats):
            for k, ms in br.profiled(mode):
                acc[k] = acc.get(k, 0.0) + ms / repeats
        out[name] = {k: round(v, 4) for k, v in sorted(acc.items(), key=lambda kv: -kv[1])}
        out[name + "_total_ms"] = round(sum(acc.values()), 3)
    cnt = np.stack([v.counters() for v in br.views]).astype(np.int64)
    out["surviving_meshlets"] = {"view": int(cnt[:, abi.CNT_DRAWS].sum()),
                                 "shadow": int(cnt[:, abi.CNT_DRAWS + 1: abi.CNT_DRAWS + 5].sum())}
    br.close()
    return out


def committed_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    runs (FETCH_SIZE, WRITE_SIZE scaled by the copy kernel's), which were
    taken on the large workload; None for another workload or kernel."""
    tfile = ROOT / "profiles" / "traffic.json"
    if workload != "large" or not tfile.exists():
        return None
    entry = json.loads(tfile.read_text()).get(kernel)
    return entry["bytes"] if entry else None


def _tuning_of(cls: int) -> dict:
    import ctypes as C

    from lighthugger_amd.lib import Tuning, lib

    t = Tuning()
    lib().lh_get_tuning_class(cls, C.byref(t))
    return t.as_dict()


def cpu_baseline(scene, cfg, uniforms, frames):
    """The oracle's whole frame on the host: its threaded form (rasters over
    draw ranges, resolve over row bands; identical outputs) on every core this
    process may use, capped at 16, timed as the median of `frames` frames after
    a warm-up; one single-thread frame beside it."""
    import statistics

    import oracle
    from oracle.compare import oracle_view

    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    sd, vd = oracle_view(scene, cfg, uniforms)
    L = oracle.lib()
    L.orc_frame_threads(sd.address, vd.address, threads)  # warm-up (page faults, caches)
    times = []
    for _ in range(frames):
        t0 = time.perf_counter()
        L.orc_frame_threads(sd.address, vd.address, threads)
        times.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    L.orc_frame(sd.address, vd.address)
    single = time.perf_counter() - t0
    cpu = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(statistics.median(times) * 1000.0, 2), "unit": "ms/frame", "cores": threads,
            "kind": "port", "single_thread_ms": round(single * 1000.0, 2),
            "sample": f"median of {frames} whole frames of the same workload (after 1 warm-up) through the "
                      f"oracle's threaded frame on {threads} threads of {cpu or 'the host'} "
                      f"(nproc={os.cpu_count()}); single_thread_ms: one frame of the scalar oracle on one thread"}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (the frame would shard "
              "sort-last over RCCL; see the design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.steps)
        return
    import torch

    from lighthugger_amd import abi
    from lighthugger_amd.frame import ViewConfig
    from lighthugger_amd.lib import Renderer
    from lighthugger_amd.scene import HostScene, make_uniforms, preset_camera

    preset = {"large": abi.PRESET_LARGE, "small": abi.PRESET_SMALL, "stress": abi.PRESET_STRESS}[args.workload]
    if args.diff_only:
        print(json.dumps(diff_bench(max(1, args.diff_steps))), flush=True)
        return
    if args.batch_profile:
        print(json.dumps(batch_profile(HostScene.preset(preset))), flush=True)
        return
    shadow = {"large": 2048, "small": 1024, "stress": 4096}[args.workload]
    t0 = time.time()
    scene = HostScene.preset(preset)
    cam = preset_camera(preset)
    cfg = ViewConfig(cam.width, cam.height, shadow, big_capacity=1 << 21, clip_capacity=1 << 18, hiz=args.hiz)
    r = Renderer(scene, cfg)
    r.set_camera(cam)
    print(f"# scene built in {time.time() - t0:.1f}s: {scene.num_instances} instances, "
          f"{scene.total_triangles} triangles", file=sys.stderr, flush=True)

    for _ in range(max(1, args.warmup)):
        r.frame()
    torch.cuda.synchronize()
    st = r.status()
    if st[abi.CNT_OVERFLOW]:
        raise RuntimeError(f"capacity overflow during warm-up: {st[abi.CNT_OVERFLOW]:#x}")
    queue_counts = {"big": [int(x) for x in st[abi.CNT_BIG: abi.CNT_BIG + 2]],
                    "clip": [int(x) for x in st[abi.CNT_CLIP: abi.CNT_CLIP + 2]]}
    if args.profile_only:
        for _ in range(args.steps):
            r.frame()
        torch.cuda.synchronize()
        print("# profile-only run done", file=sys.stderr)
        return

    # ---- timed region: exactly `steps` frames (lh_frame) ----
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        r.frame()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t_start) * 1000.0 / args.steps

    # the same frames replayed from a hipGraph (lh_frame_graph), captured
    # before its timed loop
    r.frame_graph()
    torch.cuda.synchronize()
    t_graph = time.perf_counter()
    for _ in range(args.steps):
        r.frame_graph()
    torch.cuda.synchronize()
    ms_graph = (time.perf_counter() - t_graph) * 1000.0 / args.steps
    r.close_graph()

    # the same frame with the occlusion culling switched the other way (same output)
    r.view.c.flags ^= abi.VIEW_HIZ
    r.frame()
    torch.cuda.synchronize()
    t_alt = time.perf_counter()
    for _ in range(args.steps):
        r.frame()
    torch.cuda.synchronize()
    ms_alt = (time.perf_counter() - t_alt) * 1000.0 / args.steps
    r.view.c.flags ^= abi.VIEW_HIZ
    r.frame()
    torch.cuda.synchronize()
    st_hiz = r.status()

    # ---- per-kernel device times over the same number of frames ----
    acc: dict[str, list[float]] = {}
    for _ in range(args.steps):
        for name, k_ms in r.frame_profiled():
            acc.setdefault(name, []).append(k_ms)
    # kernels launched several times per frame (per cascade) are summed per frame
    per_frame = {k: sum(v) / args.steps for k, v in acc.items()}
    per_launch = {k: float(np.mean(v)) for k, v in acc.items()}
    dominant = max(per_frame, key=per_frame.get)

    # ---- workload statistics ----
    tri_per_instance = scene.instance_triangles()
    _, _, vis0 = r.view.prefix(0)
    visible_candidates = int(tri_per_instance[vis0["index"].astype(np.int64)].sum())
    counters = r.view.counters()
    vis = r.view.visbuffer()
    covered = int((vis != np.uint64(abi.LH_VISBUFFER_CLEAR)).sum())
    sizes = meshlet_sizes(scene)

    # ---- raster statistics (one extra frame with LH_VIEW_STATISTICS) ----
    r.view.c.flags = abi.VIEW_STATISTICS
    r.frame()
    torch.cuda.synchronize()
    sc = r.view.counters()
    r.view.c.flags = 0
    frags = [int(x) for x in sc[abi.CNT_FRAGMENTS: abi.CNT_FRAGMENTS + abi.NUM_PASSES]]
    atomics = [int(x) for x in sc[abi.CNT_ATOMICS: abi.CNT_ATOMICS + abi.NUM_PASSES]]
    raster_atomics = {"fragments_per_pass": frags, "atomics_per_pass": atomics}
    for kind, lo, hi in (("view", 0, 1), ("shadow", 1, abi.NUM_PASSES)):
        kt = sum(per_frame.get(f"{k}_{kind}", 0.0)
                 for k in ("raster_meshlets", "raster_clip", "raster_big", "tile_raster", "raster_direct"))
        if kt > 0:
            raster_atomics[f"{kind}_atomics_per_s_G"] = round(sum(atomics[lo:hi]) / (kt * 1e-3) / 1e9, 2)
    if not args.no_atomic_probe:
        raster_atomics["ceiling_u64_1080p_G"] = atomic_ceiling(cfg.width - cfg.width % 8, cfg.height - cfg.height % 8,
                                                               64)
        raster_atomics["ceiling_u32_shadow_G"] = atomic_ceiling(shadow, shadow, 32)

    roofline = None
    formulas = ("raster_meshlets_view", "raster_meshlets_shadow", "tile_raster_view", "resolve")
    known = [k for k in formulas if k in per_frame]
    if dominant not in known and known:
        dominant = max(known, key=per_frame.get)  # the costliest kernel whose work is counted
    if dominant in ("tile_raster_view", "resolve"):
        if dominant == "tile_raster_view":
            algo = tile_algorithmic_bytes(scene, r.view, sizes, covered)
        else:
            algo = resolve_algorithmic_bytes(scene, r.view, covered)
        launches = max(1, round(len(acc[dominant]) / args.steps))
        algo //= launches
        achieved = algo / (per_launch[dominant] * 1e-3) / 1e9
        traffic = committed_traffic(dominant, args.workload)
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                    "algorithmic_bytes": algo, "avg_launch_ms": round(per_launch[dominant], 5),
                    "launches_per_frame": launches}
    elif dominant.startswith("raster_meshlets"):
        if dominant.endswith("_view"):
            algo = raster_algorithmic_bytes(scene, r.view.draws(0), covered, sizes)
        else:
            shadow_maps = r.view.shadow_maps()
            algo = 0
            for c in range(abi.NUM_CASCADES):
                cov_c = int((shadow_maps[c] != np.uint32(abi.LH_SHADOW_CLEAR)).sum())
                algo += raster_algorithmic_bytes(scene, r.view.draws(1 + c), cov_c, sizes) - 4 * cov_c
            algo //= abi.NUM_CASCADES
        # the main view launches twice per frame with occlusion culling (two
        # phases): bytes per launch = the frame's bytes over its launches
        launches = max(1, round(len(acc[dominant]) / args.steps))
        if dominant.endswith("_view"):
            algo //= launches
        achieved = algo / (per_launch[dominant] * 1e-3) / 1e9
        traffic = committed_traffic(dominant, args.workload)
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                    "algorithmic_bytes": algo, "avg_launch_ms": round(per_launch[dominant], 5),
                    "launches_per_frame": len(acc[dominant]) // args.steps}
    else:
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": None, "traffic": None, "avg_launch_ms": round(per_launch[dominant], 5)}

    quantized = None
    if args.quantized_steps > 0:
        quantized = variant_bench(preset, cfg, cam, args.quantized_steps, args.warmup,
                                  vertex_format=abi.VERTEX_FORMAT_QUANTIZED)
        quantized["geometry_bytes_fp32"] = int(scene.geometry.nbytes)
        torch.cuda.empty_cache()
    bc1 = None
    if args.bc1_steps > 0:
        bc1 = variant_bench(preset, cfg, cam, args.bc1_steps, args.warmup, texture_format=abi.TEXTURE_BC1)
        bc1["texel_bytes_rgba8"] = int(scene.texels.nbytes)
        torch.cuda.empty_cache()

    batched = None
    if args.batch_steps > 0 and args.workload == "large":
        del r
        torch.cuda.empty_cache()
        batched = batch_bench(scene, args.batch_steps)

    differentiable = None
    if args.diff_steps > 0 and args.workload == "large":
        torch.cuda.empty_cache()
        differentiable = diff_bench(args.diff_steps)

    cpu = None
    if args.cpu_frames > 0:
        cpu = cpu_baseline(scene, cfg, make_uniforms(cam, scene.num_instances), args.cpu_frames)

    out = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "ms_per_step_graph_replay": round(ms_graph, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+i64",
        "data": "synthetic (seeded procedural scene)",
        "gtri_per_s": round(visible_candidates / (ms * 1e-3) / 1e9, 3),
        "frames_per_s": round(1000.0 / ms, 1),
        "config": {
            "workload": f"plain frame, {args.workload}",
            "instances": scene.num_instances,
            "meshes": scene.num_meshes,
            "scene_triangles": int(scene.total_triangles),
            "visible_candidate_triangles": visible_candidates,
            "resolution": f"{cfg.width}x{cfg.height}",
            "cascades": f"4x{shadow}^2",
            "surviving_meshlets_per_pass": [int(x) for x in counters[:5]],
            "covered_pixels": covered,
            "queued_triangles_view_shadow": queue_counts,
            "tuning": {str(c): _tuning_of(c) for c in range(5)},
        },
        "kernel_ms_per_frame": {k: round(v, 4) for k, v in sorted(per_frame.items(), key=lambda kv: -kv[1])},
        "roofline": roofline,
        "raster_atomics": raster_atomics,
        "occlusion_culling": {
            "enabled": cfg.hiz, "ms_per_frame_other_setting": round(ms_alt, 4),
            "main_view_meshlets_drawn": int(st_hiz[abi.CNT_DRAWS]),
            "phase1_meshlets": int(st_hiz[abi.CNT_HIZ_PHASE1]) if cfg.hiz else None,
            "candidates": int(st_hiz[abi.CNT_HIZ_CANDIDATES]) if cfg.hiz else None,
        },
        "cpu_baseline": cpu,
        "quantized_vertices": quantized,
        "bc1_textures": bc1,
        "batched_views": batched,
        "differentiable_resolve": differentiable,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
