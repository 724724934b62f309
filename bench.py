"""Benchmark of the KNC per-block RK stage on one MI355X.

Workload: the SRHD kinetic-bomb + jet mesh, 4096 x 1024 zones in 32 uneven
radial blocks, gamma = 4/3, HLLE, RK2, PLM theta = 1.5, CFL 0.4 with the CFL
reduction run once per fold of 10 steps.  A step is one full RK2 step
(two stage launches) over every zone; the metric is FP64 zone-updates/s
(zones x steps / wall seconds), the reference's Mzps x 1e6
(src/main.rs:42-48).

Paths timed on the same initial data:
  plain    one stage launch per block per stage (KNC_LAUNCH_PER_BLOCK),
           plain c2p guess, shape autotuned like the batched paths
  batched  one persistent launch per stage over every block, the fold's
           steps replayed as a hipGraph (KNC_LAUNCH_BATCHED), plain guess
  learned  batched + the learned c2p pressure guess (BF16-MFMA network,
           weights from knc/guess_net.npz, trained by knc.train_guess),
           evaluated for every stage (learned) or once per RK step on the
           step's input state (learned_step)
The reported value is the faster learned + batched form (batched when no
trained weights exist); the other rates, the rel-L1 of every final state
against the plain path, and Newton-iteration statistics are reported beside
it.

Roofline of the headline path's stage: algorithmic HBM bytes per zone-update
are 40 B (5 FP64 fields) read + 40 B written on stage 1 and 80 B read + 40 B
written on stage 2 (fused RK average), i.e. 100 B per zone per stage on
average; achieved = 100 B x zones / the stage's average device time (k_stage
plus its share of the learned-guess pass), measured with HIP events on the
context's stream over the timed region.  The plain-guess batched leg's
k_stage alone gets the same figure as plain_stage_roofline.  The FP64 VALU
roofline (valu_roofline) is the kernel's actual bound.

The default run is 2000 timed steps per GPU path (the benchmark's length).

CPU baseline: the C++ oracle (oracle/) on the host cores, a bounded sample
of the same mesh (100 RK2 steps after one warm-up step by default, about
10-20 s on 16 threads), reported as zone-updates/s.

The batched paths run the library's autotuner (knc_autotune) on the mesh
before timing; the chosen shapes are reported.  A short FP64 FMA probe
(knc_fp64_fma_rate, fma_probe) measures the VALU issue peak the stage is
compared with, and the rate one dependent chain leaves at two waves per SIMD.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "app-kilonova_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
BYTES_PER_ZONE_STAGE = 100.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(setup, launch_mode, steps, warmup, fold, net=None, tune=False, guess=None, **shape):
    import knc
    if guess is None:
        guess = knc.GUESS_LEARNED if net is not None else knc.GUESS_PLAIN
    ctx = knc.context_from_setup(setup, launch_mode=launch_mode, use_graph=True, c2p_guess=guess, **shape)
    if net is not None:
        net.attach(ctx)
    ctx.tuned = ctx.autotune() if tune else None
    if ctx.tuned:
        log(f"bench: autotuned shape {ctx.tuned['stage_threads']} columns x {ctx.tuned['rows_per_item']} rows, "
            f"stage form {ctx.tuned['stage_passes']} of {ctx.tuned['candidates']} candidates "
            f"({ctx.tuned['ms_per_stage']:.4f} ms vs default {ctx.tuned['default_ms_per_stage']:.4f} ms)")
    ctx.set_timing(True)
    done = 0
    while done < warmup:
        f = min(fold, warmup - done)
        ctx.advance(f)
        done += f
    ctx.synchronize()
    stage_ms = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        f = min(fold, steps - done)
        ctx.advance(f)
        ms, _ = ctx.last_stage_time()
        stage_ms.append((ms, f))
        done += f
    ctx.synchronize()
    wall = time.perf_counter() - t0
    avg_stage_ms = sum(m * f for m, f in stage_ms) / sum(f for _, f in stage_ms)
    return ctx, wall, avg_stage_ms


def cpu_model():
    """The host CPU's model name (lscpu's "Model name", from /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(setup, nsteps, nthreads):
    import oracle  # noqa: F401  (the checker / CPU baseline only)
    from util import oracle_from_setup
    sim = oracle_from_setup(setup, nthreads)
    dt = sim.time_step()
    sim.advance(dt, 1)  # warm-up step
    t0 = time.perf_counter()
    dt = sim.time_step()
    sim.advance(dt, nsteps)
    wall = time.perf_counter() - t0
    return setup.total_zones * nsteps / wall, wall


def newton_stats(setup, fold, net, steps=20):
    """Newton iteration histogram and fallback rate of the plain, learned and
    previous-pressure guesses over a short instrumented run (not timed).  The two-pass stage
    form is used so that every zone is counted once per stage."""
    import knc
    out = {}
    for name, n, guess in (("plain", None, knc.GUESS_PLAIN), ("learned", net, knc.GUESS_LEARNED),
                           ("learned_step", net, knc.GUESS_LEARNED_STEP), ("learned_fold", net, knc.GUESS_LEARNED_FOLD),
                           ("previous", None, knc.GUESS_PREVIOUS)):
        if name.startswith("learned") and n is None:
            continue
        ctx = knc.context_from_setup(setup, launch_mode=knc.BATCHED, c2p_guess=guess, collect_stats=True,
                                     stage_passes=2)
        if n is not None:
            n.attach(ctx)
        for _ in range(steps // fold):
            ctx.advance(fold)
        st = ctx.c2p_stats()
        h = st["hist"]
        z = max(1, int(h.sum()))
        out[name] = {"mean_iterations": float((h * np.arange(h.size)).sum() / z),
                     "histogram": {int(k): int(v) for k, v in enumerate(h) if v},
                     "fallback_rate": st["fallbacks"] / z, "floors": st["floors"], "zones": z}
    return out


def read_traffic(head):
    """HBM bytes per stage of the plain-guess path (head None) or of a
    learned-guess path ("learned", "learned_step": the stage kernel plus its
    share of the guess pass) from the committed rocprofv3 PMC summary, or
    None."""
    path = os.path.join(ROOT, "profiles", "k_stage_bomb_jet_pmc.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_stage_launch" + (f"_{head}" if head else "")])
    except (KeyError, ValueError):
        return None


VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4   # CUs x SIMDs x clock / cycles per wave64 FP64 instruction


def fma_probe():
    """FP64 FMA issue rates measured on the device (knc_fp64_fma_rate), in G
    wave-instructions/s: eight independent chains per lane at eight waves per
    SIMD (the issue peak as measured), and four chains and one dependent
    chain at two waves per SIMD, the stage kernel's occupancy."""
    import knc
    out = {"unit": "G wave-instructions/s", "note": "best of 3 launches of 16384 FMAs per chain"}
    for name, chains, waves in (("8_chains_8_waves", 8, 8), ("4_chains_2_waves", 4, 2), ("1_chain_2_waves", 1, 2)):
        rate = max(knc.fp64_fma_rate(chains, waves, 16384)[0] for _ in range(3))
        out[name] = rate / 64 / 1e9
    return out


def valu_roofline(zones, stage_ms, probe=None):
    """VALU issue rate of the plain-guess stage kernel: instructions per zone
    from the committed SQ_INSTS_VALU summary x zones / measured stage time,
    against the MI355X issue peak (and the measured FMA issue peak), or None."""
    path = os.path.join(ROOT, "profiles", "k_stage_bomb_jet_sq.txt")
    if not os.path.exists(path):
        return None
    per_zone = None
    for line in open(path):
        if line.startswith("valu_insts_per_zone"):
            per_zone = float(line.split("=", 1)[1])
    if per_zone is None:
        return None
    achieved = per_zone * zones / 64 / (stage_ms * 1e-3)
    out = {"kernel": "k_stage, plain guess", "bound": "fp64-valu", "valu_insts_per_zone": per_zone,
           "achieved": achieved / 1e9, "peak": VALU_PEAK_WAVE_INSTS / 1e9, "unit": "G wave-instructions/s",
           "frac": achieved / VALU_PEAK_WAVE_INSTS}
    if probe:
        out["measured_fma_peak"] = probe["8_chains_8_waves"]
        out["frac_of_measured_peak"] = achieved / 1e9 / probe["8_chains_8_waves"]
    return out


# FP64 operations of one Newton iteration of the c2p as restated (add, sub,
# mul, div, sqrt each 1): Q = E + p, den = Q^2 - S^2 (2), sqrt, p Q,
# F = (gamma-1)(den - D sq - pQ) - pQ (5), dp = F den / (gamma p S^2 - den Q) (6),
# p - dp, the convergence test 1e-12 (E + p) (2).
NEWTON_FLOPS = 19
FP64_VECTOR_PEAK = 78.6e12


def c2p_roofline(zones, stage_ms, mean_iterations):
    """Time the c2p Newton solves of one stage would take at the FP64 vector
    peak, as a fraction of the measured stage time."""
    flops = zones * mean_iterations * NEWTON_FLOPS
    t = flops / FP64_VECTOR_PEAK
    return {"flops_per_iteration": NEWTON_FLOPS, "mean_iterations": mean_iterations, "gflop_per_stage": flops / 1e9,
            "ms_at_fp64_peak": 1e3 * t, "frac_of_stage": 1e3 * t / stage_ms, "peak_tflops": FP64_VECTOR_PEAK / 1e12}


def io_rates(ctx, zones):
    """The reference's output files of the bench state, once, outside the
    timed region: knc_write_checkpoint (state streamed from HBM and encoded
    on the host), knc_write_products (device c2p, then the same), and the
    checkpoint reader (parse and check, then decode every block)."""
    import tempfile

    from knc import app as kapp
    from knc import capi, cbor
    cfg = cbor.encode({"workload": "bomb_jet"})
    tasks = cbor.encode(kapp.Tasks(0.0).to_serde())
    out = {"state_MB": 40.0 * zones / 1e6}
    with tempfile.TemporaryDirectory() as d:
        for name, write in (("checkpoint", lambda p: ctx.write_checkpoint(p, cfg, tasks, kapp.VERSION)),
                            ("products", lambda p: ctx.write_products(p, cfg, kapp.VERSION))):
            path = os.path.join(d, name + ".cbor")
            ctx.synchronize()
            t0 = time.perf_counter()
            write(path)
            dt = time.perf_counter() - t0
            size = os.path.getsize(path) / 1e6
            out[name] = {"write_s": dt, "file_MB": size, "file_MB_per_s": size / dt}
        path = os.path.join(d, "checkpoint.cbor")
        t0 = time.perf_counter()
        with capi.Checkpoint(path) as ck:
            t1 = time.perf_counter()
            for k in range(len(ck.blocks)):
                ck.block_data(k)
            t2 = time.perf_counter()
        out["checkpoint"].update({"open_s": t1 - t0, "decode_s": t2 - t1,
                                  "read_MB_per_s": out["checkpoint"]["file_MB"] / (t2 - t0)})
        from knc import records
        t0 = time.perf_counter()
        records.read_products(os.path.join(d, "products.cbor"))
        out["products"]["numpy_read_s"] = time.perf_counter() - t0
    log(f"bench: io checkpoint {out['checkpoint']['file_MB']:.0f} MB written in {out['checkpoint']['write_s']:.3f} s,"
        f" read in {out['checkpoint']['open_s'] + out['checkpoint']['decode_s']:.3f} s; products "
        f"{out['products']['file_MB']:.0f} MB in {out['products']['write_s']:.3f} s")
    return out


def copy_probe(mib=512, repeats=5):
    """Device-to-device copies of a known size, so that FETCH_SIZE and
    WRITE_SIZE per dispatch can be scaled to bytes."""
    import torch
    src = torch.empty(mib << 20, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(repeats):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"bench: copy probe, {repeats} x {mib} MiB device-to-device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--fold", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--nr", type=int, default=4096)
    ap.add_argument("--nq", type=int, default=1024)
    ap.add_argument("--threads", type=int, default=0, help="stage workgroup size (0: library choice)")
    ap.add_argument("--rows", type=int, default=0, help="radial rows per work item (0: library choice)")
    ap.add_argument("--passes", type=int, default=0,
                    help="stage form: 1 fused, 2 two-pass, 3 warp-specialised (0: library choice)")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-previous", action="store_true", help="skip the previous-pressure guess path")
    ap.add_argument("--no-stats", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (library default shape)")
    ap.add_argument("--no-io", action="store_true", help="skip the checkpoint/products file measurement")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run five 512 MiB device-to-device copies (calibration of the HBM traffic counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs on a single MI355X; sharding the radial block "
              "chain across GPUs (halo rows per cut, a min-reduction of dt per fold) is not implemented",
              file=sys.stderr)
        sys.exit(2)

    import knc
    setup = knc.setups.bomb_jet(nr=args.nr, nq=args.nq)
    zones = setup.total_zones
    log(f"bench: bomb_jet {args.nr}x{args.nq} zones={zones} blocks={len(setup.blocks)} "
        f"steps={args.steps} warmup={args.warmup} fold={args.fold}")

    shape = dict(stage_threads=args.threads, rows_per_item=args.rows, stage_passes=args.passes)
    tune = not args.no_tune and not (args.threads or args.rows or args.passes)
    ctx_b, wall_b, stage_ms_b = run_gpu(setup, knc.BATCHED, args.steps, args.warmup, args.fold, tune=tune, **shape)
    rate_b = zones * args.steps / wall_b
    log(f"bench: batched {rate_b / 1e9:.3f} Gzu/s, stage {stage_ms_b:.4f} ms")

    net = None
    if not args.no_learned:
        from knc import train_guess
        try:
            net = train_guess.load()
        except RuntimeError as e:
            log(f"bench: learned path skipped: {e}")
    learned = {}
    ctx_l = {}
    if net is not None:
        for name, guess in (("learned", knc.GUESS_LEARNED), ("learned_step", knc.GUESS_LEARNED_STEP),
                            ("learned_fold", knc.GUESS_LEARNED_FOLD)):
            ctx, wall, stage_ms = run_gpu(setup, knc.BATCHED, args.steps, args.warmup, args.fold, net=net,
                                          tune=tune, guess=guess, **shape)
            rate = zones * args.steps / wall
            log(f"bench: {name}+batched {rate / 1e9:.3f} Gzu/s, stage {stage_ms:.4f} ms")
            learned[name] = {"value": rate, "ms_per_step": 1e3 * wall / args.steps, "stage_ms": stage_ms}
            ctx_l[name] = ctx

    # the previous-pressure guess: every stage starts Newton from the
    # pressure its zone's previous stage recovered (no network, no extra pass)
    prev = None
    if not args.no_previous:
        ctx_v, wall_v, stage_ms_v = run_gpu(setup, knc.BATCHED, args.steps, args.warmup, args.fold, tune=tune,
                                            guess=knc.GUESS_PREVIOUS, **shape)
        rate_v = zones * args.steps / wall_v
        log(f"bench: previous+batched {rate_v / 1e9:.3f} Gzu/s, stage {stage_ms_v:.4f} ms")
        prev = {"value": rate_v, "ms_per_step": 1e3 * wall_v / args.steps, "stage_ms": stage_ms_v}
        ctx_l["previous"] = ctx_v

    plain = None
    parity = {}
    if not args.no_plain:
        # the per-block path is tuned too, so that each block's launch fills
        # the device as well as its size allows
        ctx_p, wall_p, stage_ms_p = run_gpu(setup, knc.PER_BLOCK, args.steps, args.warmup, args.fold, tune=tune,
                                            **shape)
        rate_p = zones * args.steps / wall_p
        log(f"bench: plain {rate_p / 1e9:.3f} Gzu/s, stage {stage_ms_p:.4f} ms")
        from util import gpu_state, rel_l1  # noqa: E402
        ref = gpu_state(ctx_p)
        parity["batched_vs_plain"] = rel_l1(gpu_state(ctx_b), ref)
        for name, ctx in ctx_l.items():
            parity[f"{name}_batched_vs_plain"] = rel_l1(gpu_state(ctx), ref)
        plain = {"value": rate_p, "ms_per_step": 1e3 * wall_p / args.steps, "stage_ms": stage_ms_p}

    newton = newton_stats(setup, args.fold, net) if not args.no_stats else None
    io = io_rates(ctx_b, zones) if not args.no_io else None
    probe = fma_probe()
    log(f"bench: FP64 FMA issue {probe['8_chains_8_waves']:.1f} G wave-instructions/s (8 chains, 8 waves/SIMD), "
        f"{probe['1_chain_2_waves']:.1f} for one dependent chain at 2 waves/SIMD")

    # headline: the faster learned-guess form (batched plain guess without weights)
    head = max(learned, key=lambda k: learned[k]["value"]) if learned else None
    head_rate, head_stage = (learned[head]["value"], learned[head]["stage_ms"]) if head else (rate_b, stage_ms_b)
    # Roofline of the headline path's stage: k_stage and, for the learned
    # paths, its share of the guess pass (k_guess once per RK step for
    # learned_step, every stage for learned), timed together with HIP events
    # on the context's stream around each fold's graph.  The plain-guess
    # batched leg's stage, k_stage alone, is reported beside it.
    achieved = BYTES_PER_ZONE_STAGE * zones / (stage_ms_b * 1e-3) / 1e9
    achieved_head = BYTES_PER_ZONE_STAGE * zones / (head_stage * 1e-3) / 1e9
    head_kernel = {"learned": "k_stage (guess plane) + k_guess every stage, batched",
                   "learned_step": "k_stage (guess plane) + half a k_guess pass (once per RK2 step), batched",
                   "learned_fold": f"k_stage (guess plane) + one k_guess pass per fold ({2 * args.fold} stages), "
                                   "batched",
                   None: "k_stage (plain guess), batched"}[head]
    cpu = None
    if not args.no_cpu:
        # the host threads this process may run on, as the reference sizes its
        # runtime from the machine's cores (src/app.rs:148-153)
        ncores = len(os.sched_getaffinity(0))
        rate_c, wall_c = cpu_baseline(setup, args.cpu_steps, ncores)
        # the reference's default worker count is twice the core count (src/app.rs:148-153)
        steps2 = max(1, args.cpu_steps // 2)
        rate_2x, wall_2x = cpu_baseline(setup, steps2, 2 * ncores)
        cpu = {"value": rate_c, "unit": "zone-updates/s", "cores": ncores, "kind": "port", "cpu_model": cpu_model(),
               "sample": f"oracle (C++ restatement), bomb_jet {args.nr}x{args.nq}, {args.cpu_steps} RK2 steps "
                         f"after 1 warm-up step, {wall_c:.1f} s",
               "threads_2x_cores": {"value": rate_2x, "threads": 2 * ncores,
                                    "sample": f"{steps2} RK2 steps, {wall_2x:.1f} s"}}
        log(f"bench: cpu {rate_c / 1e6:.3f} Mzu/s on {ncores} threads, {rate_2x / 1e6:.3f} Mzu/s on "
            f"{2 * ncores} threads")

    out = {
        "metric": "FP64 zone-updates/s (RK2 step of the per-block stage)",
        "value": head_rate,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * zones / head_rate,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded)",
        "config": {"workload": f"bomb_jet SRHD kinetic bomb + jet, {args.nr}x{args.nq} zones, "
                               f"{len(setup.blocks)} uneven blocks, HLLE, RK2, fold {args.fold}",
                   "path": ({"learned": "learned c2p guess evaluated for every stage",
                             "learned_step": "learned c2p guess evaluated once per RK step (first stage's input)",
                             "learned_fold": "learned c2p guess evaluated once per fold (first stage's input)"}
                            [head] + " + batched persistent stage launches, hipGraph per fold"
                            if head else "batched persistent stage launches, hipGraph per fold, plain c2p guess"),
                   "zones": zones},
        "roofline": {"bound": "hbm", "achieved": achieved_head, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_head / HBM_PEAK_GBS, "traffic": read_traffic(head),
                     "kernel": head_kernel, "kernel_ms": head_stage},
        "plain_stage_roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(None),
                                 "kernel": "k_stage (plain guess), batched", "kernel_ms": stage_ms_b},
        "valu_roofline": valu_roofline(zones, stage_ms_b, probe),
        "fma_probe": probe,
        "c2p_fp64_roofline": (c2p_roofline(zones, head_stage, newton[head or "plain"]["mean_iterations"])
                              if newton else None),
        "autotuned": {"batched": ctx_b.tuned, **{k: c.tuned for k, c in ctx_l.items()},
                      "plain_per_block": ctx_p.tuned if plain else None},
        "cpu_baseline": cpu,
        "paths": {"plain_per_block": plain,
                  "batched": {"value": rate_b, "ms_per_step": 1e3 * wall_b / args.steps, "stage_ms": stage_ms_b},
                  **{f"{k}_batched": v for k, v in learned.items()},
                  **({"previous_batched": prev} if prev else {})},
        "speedup_vs_plain": ({"batched": rate_b / plain["value"],
                              **{f"{k}_batched": v["value"] / plain["value"] for k, v in learned.items()},
                              **({"previous_batched": prev["value"] / plain["value"]} if prev else {})}
                             if plain else None),
        "rel_l1_vs_plain": parity or None,
        "newton": newton,
        "io": io,
        "stage_ms": head_stage,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
