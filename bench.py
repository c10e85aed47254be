#!/usr/bin/env python3
"""FAC Poisson benchmark on one MI355X.

A step is one CellPoissonFACSolver::solveSystem from u = 0 to a relative
residual of 1e-10 (RMS residual <= 1e-10 * RMS f, FP64) on the 3-D deck:
unit cube, 128^3 base in 64^3 patches, three levels at ratio 2 refined where
|grad u| exceeds the 70th percentile, boxes of 16^3..64^3; D = 1,
u = sin(pi x) sin(pi y) sin(pi z), Dirichlet 0; red-black GS 2+2 sweeps,
coarsest level red-black (500 sweeps, tolerance 1e-12), Ewing coarse-fine
fluxes, conservative linear prolongation.

Prints one JSON line: time to solution (value, s), smoother throughput, the
roofline of the smoother kernel with the largest share of the time, measured
with HIP events on the solver's stream over one more (untimed) solve, and the
CPU baseline (the C oracle, a
bounded sample of FAC cycles on the same deck, projected to the cycle count
the GPU needed).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

PARAMS = dict(max_cycles=200, residual_tol=1e-300, relative_residual_tol=1e-10,
              num_pre_sweeps=2, num_post_sweeps=2, coarse_solver_choice="redblack",
              coarse_solver_tolerance=1e-12, coarse_solver_max_iterations=500,
              cf_discretization="Ewing", prolongation_method="CONSERVATIVE_LINEAR_REFINE")


def build_problem(fac, decks, deck, params=None):
    h = fac.PatchHierarchy(deck["dim"], deck["dlo"], deck["dhi"], [0.0] * 3, [1.0] * 3)
    for boxes in deck["levels"]:
        h.add_level(deck["ratio"], boxes)
    h.finalize()
    u = fac.PatchData(h, 1)
    f = fac.PatchData(h, 0)
    for ln, boxes in enumerate(deck["levels"]):
        for p, box in enumerate(boxes):
            f.put(ln, p, decks.sine_rhs(deck, ln, box))
    s = fac.CellPoissonFACSolver(h, **(params or PARAMS))
    s.set_d_constant(1.0)
    s.set_dirichlet(0.0)
    s.initialize_solver_state(u, f)
    return h, u, f, s


MG_PARAMS = dict(PARAMS, coarse_solver_choice="mg", coarse_solver_tolerance=1e-8,
                 coarse_solver_max_iterations=10)


def max_error(decks, deck, u):
    err = 0.0
    for ln, boxes in enumerate(deck["levels"]):
        for p, box in enumerate(boxes):
            shp = tuple(h - l + 3 for l, h in zip(*box))
            uu = u.get(ln, p).reshape(shp, order="F")[1:-1, 1:-1, 1:-1]
            err = max(err, float(np.max(np.abs(uu - decks.sine_exact(deck, ln, box)))))
    return err


def cpu_baseline(decks, deck, cycles_gpu, ncycles, threads, mg=False):
    from oracle.oracle import OracleFAC
    o = OracleFAC(3, deck["levels"], deck["ratio"], deck["dlo"], deck["dhi"], [0.0] * 3,
                  [1.0] * 3)
    o.set_poisson(False, 1.0, 0, 0.0)
    o.set_bc([1.0] * 6, [0.0] * 6, [0.0] * 6)
    P = MG_PARAMS if mg else PARAMS
    o.set_params(max_cycles=ncycles, residual_tol=1e-300, relative_residual_tol=1e-10,
                 pre=2, post=2, coarse_max_iterations=P["coarse_solver_max_iterations"],
                 coarse_tol=P["coarse_solver_tolerance"], ewing=True, nthreads=threads)
    o.set_coarse_mg(mg)
    for ln, boxes in enumerate(deck["levels"]):
        for p, box in enumerate(boxes):
            o.put("f", ln, p, decks.sine_rhs(deck, ln, box))
    conv, it, res, hist = o.solve()
    secs = o.last_solve_seconds()
    o.close()
    per_cycle = secs / max(it, 1)
    return dict(value=per_cycle * cycles_gpu, unit="s", cores=threads, kind="port",
                sample="%d FAC cycles of the same deck with the C oracle on %d host threads "
                       "(%.2f s, %.3f s/cycle), projected to the %d cycles the GPU solve took"
                       % (it, threads, secs, per_cycle, cycles_gpu))


def setup_c2(fac, decks, torch, params, seed=0):
    """C2: 4 levels, ~2100 patches with sides from U{8..48}, D = exp(GRF)
    (correlation length 0.1, sigma 1) on the faces, f = eight Gaussian
    sources of width 0.02; generated on the GPU and copied device to device."""
    deck = decks.deck_3d_c2(seed)
    h = fac.PatchHierarchy(3, deck["dlo"], deck["dhi"], [0.0] * 3, [1.0] * 3)
    for boxes in deck["levels"]:
        h.add_level(deck["ratio"], boxes)
    h.finalize()
    u = fac.PatchData(h, 1)
    f = fac.PatchData(h, 0)
    D = fac.PatchData(h, 0, "side")
    modes = decks.grf_modes(seed)
    kw = dict(dtype=torch.float64, device="cuda")
    for ln, boxes in enumerate(deck["levels"]):
        for p, box in enumerate(boxes):
            fv = decks.gaussian_sources(deck, decks.cell_points(deck, ln, box, xp=torch, **kw),
                                        xp=torch)
            f.put_device(ln, p, fv.permute(2, 1, 0).contiguous())
            for a in range(3):
                g = decks.grf_eval(decks.face_points(deck, ln, box, a, xp=torch, **kw), modes,
                                   xp=torch)
                D.put_device(ln, p, torch.exp(g).permute(2, 1, 0).contiguous(), axis=a)
    s = fac.CellPoissonFACSolver(h, **params)
    s.set_d_patch_data(D)
    s.set_dirichlet(0.0)
    s.initialize_solver_state(u, f)
    return deck, h, u, f, D, s


def measured_traffic(tile):
    """L2-miss bytes per launch of the fused sweep with this tile shape on the
    C1 coarsest level, from the committed counter summary (None if absent)."""
    path = os.path.join(ROOT, "profiles", "k_box_rbgs_c1_hbm_traffic.txt")
    if tile is None or not os.path.exists(path):
        return None
    key = "x".join(map(str, tile))
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == key:
            return int(f[3])
    return None


def level_errors(deck, ua, ub):
    """Per-level RMS and max |ua - ub| over interior cells (ub: FP64 reference)."""
    out = []
    for ln, boxes in enumerate(deck["levels"]):
        sq, n, mx = 0.0, 0, 0.0
        for p, box in enumerate(boxes):
            shp = tuple(h - l + 3 for l, h in zip(*box))
            a = ua.get(ln, p).reshape(shp, order="F")[1:-1, 1:-1, 1:-1]
            b = ub.get(ln, p).reshape(shp, order="F")[1:-1, 1:-1, 1:-1]
            d = a - b
            sq += float(np.sum(d * d))
            n += d.size
            mx = max(mx, float(np.max(np.abs(d))))
        out.append(dict(level=ln, l2=float(np.sqrt(sq / max(n, 1))), max=mx))
    return out


def time_c2(fac, decks, torch, params, bits=64):
    deck, h, u, f, D, s = setup_c2(fac, decks, torch, params)
    s.set_coarse_precision(bits)
    u.fill(0.0)
    s.solve_system(u, f)  # warm-up (graph capture, autotuning)
    u.fill(0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ok = s.solve_system(u, f)
    s.synchronize()
    dt = time.perf_counter() - t0
    out = dict(s_per_solve=round(dt, 4), cycles=s.number_of_iterations(), converged=bool(ok),
               final_residual=float(s.history()[-1]))
    if params.get("coarse_solver_choice") == "mg":
        out["last_coarse_vcycles"] = s.coarse_mg_stats()[0]
    return deck, out, (h, u, f, D, s)


MFMA_BF16_PEAK_TFLOPS = 2500.0
# multiply-adds per cell of the network as specified (2 -> 16 -> 16 -> 1
# channels, 27 taps); the kernels pad every layer to 16 x 16 channels
LEARNED_MACS_PER_CELL = 27 * (2 * 16 + 16 * 16 + 16 * 1)
LEARNED_PADDED_MACS_PER_CELL = 27 * 16 * 16 * 3


def learned_section(torch, seeds, steps, batch, log=print):
    """C4 (learned smoother): pairs generated by the FAC solver on seeded
    variable-coefficient problems (one-level 64^3 hierarchies of 16^3 patches,
    residuals after 1, 2 and 3 FAC cycles, targets the FP64 block solutions),
    training on the GPU by the library's own forward/backward kernels, the
    matrix-core rate of the forward convolutions, and the gain on unseen seeds."""
    from samrai_amd import learned as SL
    t0 = time.perf_counter()
    train = [SL.fac_training_pairs(torch, s, cycles=(1, 2, 3)) for s in range(seeds)]
    test = [SL.fac_training_pairs(torch, 1000 + s, cycles=(1, 2, 3)) for s in range(2)]
    cat = {k: torch.cat([p[k] for p in train]).contiguous() for k in ("r", "logd", "e")}
    tst = {k: torch.cat([p[k] for p in test]).contiguous()
           for k in ("r", "logd", "e", "dxf", "dyf", "dzf")}
    h = train[0]["h"]
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    log("bench: learned: %d training blocks from %d seeds, %d test blocks (%.1f s)" %
        (cat["r"].shape[0], seeds, tst["r"].shape[0], gen_s))
    nmax = max(batch, tst["r"].shape[0], 512)
    net = SL.LearnedSmoother(torch, nmax, seed=0)

    def quality():
        e = net.correction(tst["r"], tst["logd"], h)
        ne = torch.linalg.vector_norm(tst["e"])
        both = SL.block_sor(torch, tst["r"], tst["dxf"], tst["dyf"], tst["dzf"], h, 1, 1.0, e0=e)
        return (torch.linalg.vector_norm(e - tst["e"]) / ne).item(), \
            (torch.linalg.vector_norm(both - tst["e"]) / ne).item()

    err0, _ = quality()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    losses = net.train(cat["r"], cat["logd"], cat["e"], h, steps=steps, batch=batch, lr=2e-3,
                       log_every=max(1, steps // 10))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t1
    err1, err_both = quality()
    ne = torch.linalg.vector_norm(tst["e"])
    gs = [(torch.linalg.vector_norm(SL.block_sor(torch, tst["r"], tst["dxf"], tst["dyf"],
                                                 tst["dzf"], h, k, 1.0) - tst["e"]) / ne).item()
          for k in (1, 2, 4)]
    # matrix-core rate of the forward pass (three conv launches) on the test blocks
    rb, lb = tst["r"], tst["logd"]
    nb = rb.shape[0]
    net.forward(rb, lb, h)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 50
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(reps):
        net.forward(rb, lb, h)
    ev1.record()
    torch.cuda.synchronize()
    fwd_ms = ev0.elapsed_time(ev1) / reps
    cells = nb * 4096
    alg_tflops = 2.0 * LEARNED_MACS_PER_CELL * cells / (fwd_ms * 1e-3) / 1e12
    padded_tflops = 2.0 * LEARNED_PADDED_MACS_PER_CELL * cells / (fwd_ms * 1e-3) / 1e12
    # one red-black sweep of the same blocks (FP64), for the cost comparison
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(reps):
        SL.block_sor(torch, rb, tst["dxf"], tst["dyf"], tst["dzf"], h, 1, 1.0)
    ev1.record()
    torch.cuda.synchronize()
    gs_ms = ev0.elapsed_time(ev1) / reps
    # smoothing a level of 16^3 patches (unseen seed, 64^3 cells): red-black
    # (2 sweeps) per iteration, with and without the learned correction first;
    # the FP64 RMS residual judges both
    P = SL.BlockLevelProblem(torch, 2000)
    sv = P.solver()
    level = {}
    for learned in (False, True):
        P.u.fill(0.0)
        hist = []
        t_it = 0.0
        for it in range(8):
            sv.synchronize()
            t2 = time.perf_counter()
            if learned:
                sv.residual(P.r, P.u, P.f, 0, False)
                SL.learned_correct(sv, net, P.u, P.r, P.D, 0)
            sv.smooth(P.u, P.f, 0, 2)
            sv.synchronize()
            t_it += time.perf_counter() - t2
            sv.residual(P.r, P.u, P.f, 0, False)
            hist.append(sv.rms_norm(P.r))
        level["learned_then_rbgs2" if learned else "rbgs2"] = dict(
            rms_residual=[float("%.4e" % v) for v in hist], ms_per_iteration=round(t_it / 8 * 1e3, 3))
    sv.close()
    P.close()
    # whole FAC solves to 1e-10 on a two-level problem (64^3 + 128^3 in 16^3
    # patches, unseen seed): plain red-black smoothing against the learned
    # correction before it (omega 1 and 0.5), multigrid coarsest-level solver
    P = SL.BlockLevelProblem(torch, 3000, n=128, two_level=True)
    sv = P.solver(**dict(MG_PARAMS, max_cycles=100))
    fac_solves = {}
    u_ref = None
    for name, om, phases in (("rbgs", None, 3), ("learned_omega_1", 1.0, 3),
                             ("learned_omega_0.5", 0.5, 3), ("learned_pre_only", 1.0, 1),
                             ("learned_post_only", 1.0, 2)):
        SL.set_learned_smoother(sv, None if om is None else net, om or 1.0)
        SL.set_learned_phases(sv, phases)
        P.u.fill(0.0)
        sv.solve_system(P.u, P.f)  # graph capture
        P.u.fill(0.0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        ok = sv.solve_system(P.u, P.f)
        sv.synchronize()
        dt = time.perf_counter() - t2
        uf = torch.stack([torch.as_tensor(P.u.get(1, p)) for p in range(len(P.boxes))])
        if u_ref is None:
            u_ref = uf
        fac_solves[name] = dict(s_per_solve=round(dt, 4), cycles=sv.number_of_iterations(),
                                converged=bool(ok), final_residual=float(sv.history()[-1]),
                                learned_still_active=SL.learned_active(sv),
                                max_diff_vs_rbgs=float((uf - u_ref).abs().max()))
    SL.set_learned_smoother(sv, None)
    sv.close()
    P.close()
    log("bench: learned FAC solves %s" % json.dumps(fac_solves))
    out = dict(
        fac_solves=dict(workload="two levels, 64^3 + 128^3 in 16^3 patches (576 patches), "
                                 "D = exp(GRF), 8 Gaussian sources, unseen seed; RBGS 2+2, mg "
                                 "coarse, rel tol 1e-10", **fac_solves),
        level_smoothing=dict(workload="one-level 64^3 problem of 16^3 patches (unseen seed), "
                                      "8 iterations from u = 0", **level),
        workload="pairs from FAC iterates (1-3 cycles, 2 coarse RBGS sweeps per cycle) on "
                 "one-level 64^3 hierarchies of 16^3 patches, D = exp(GRF corr 0.1 sigma 1), "
                 "8 Gaussian sources; target = FP64 solution of the block error equation "
                 "(zero Dirichlet on the block faces, 400 SOR sweeps)",
        network="3x3x3 conv 2->16->16->1, ReLU, BF16 weights/activations, FP32 accumulation "
                "(v_mfma_f32_16x16x16_bf16), Adam on FP32 master weights",
        train_blocks=int(cat["r"].shape[0]), train_seeds=seeds, test_blocks=int(tst["r"].shape[0]),
        pair_generation_s=round(gen_s, 2), train_steps=steps, batch_blocks=batch,
        train_s=round(train_s, 3),
        train_cells_per_s=round(steps * batch * 4096 / train_s / 1e6, 1),
        train_cells_per_s_unit="Mcell/s (forward + backward + Adam)",
        loss_first_last=[losses[0][1], losses[-1][1]],
        test_block_error_ratio=dict(untrained=round(err0, 4), trained=round(err1, 4),
                                    trained_then_1_rbgs=round(err_both, 4),
                                    rbgs_1=round(gs[0], 4), rbgs_2=round(gs[1], 4),
                                    rbgs_4=round(gs[2], 4)),
        forward_ms=round(fwd_ms, 4), forward_blocks=nb, rbgs_sweep_ms=round(gs_ms, 4),
        roofline=dict(bound="mfma", achieved=round(alg_tflops, 2), peak=MFMA_BF16_PEAK_TFLOPS,
                      unit="TFLOP/s", frac=round(alg_tflops / MFMA_BF16_PEAK_TFLOPS, 5),
                      padded_tflops=round(padded_tflops, 2),
                      kernel="k_conv (3 launches, forward)",
                      flops_per_cell=2 * LEARNED_MACS_PER_CELL))
    net.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cycles", type=int, default=2)
    ap.add_argument("--smoother-sweeps", type=int, default=200)
    ap.add_argument("--no-c2", action="store_true",
                    help="skip the variable-coefficient 4-level deck (C2) section")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-smoother (C4) section")
    ap.add_argument("--learned-seeds", type=int, default=50)
    ap.add_argument("--learned-steps", type=int, default=1500)
    ap.add_argument("--learned-batch", type=int, default=64)
    ap.add_argument("--mode", choices=["solve", "smoother", "profile", "learned"],
                    default="solve",
                    help="smoother: only a few smoother launches and a 256 MiB copy of known "
                         "size (for hardware-counter passes); profile: only the solves of "
                         "--deck with --coarse (for kernel traces); neither prints the "
                         "benchmark line")
    ap.add_argument("--deck", choices=["c1", "c2"], default="c1",
                    help="deck of --mode profile")
    ap.add_argument("--coarse", choices=["redblack", "mg"], default="redblack",
                    help="coarsest-level solver of --mode profile")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from samrai_amd import decks, fac

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.mode == "learned":
        log("bench: learned %s" % json.dumps(learned_section(
            torch, args.learned_seeds, args.learned_steps, args.learned_batch, log)))
        return
    if args.mode == "profile":
        P = MG_PARAMS if args.coarse == "mg" else PARAMS
        if args.deck == "c2":
            deck, h, u, f, D, s = setup_c2(fac, decks, torch, P)
        else:
            deck = decks.deck_3d_c1()
            h, u, f, s = build_problem(fac, decks, deck, P)
        for k in range(args.warmup + args.steps):
            u.fill(0.0)
            t0 = time.perf_counter()
            s.solve_system(u, f)
            s.synchronize()
            print("bench: profile %s %s solve %d: %d cycles, %.4f s" %
                  (args.deck, args.coarse, k, s.number_of_iterations(),
                   time.perf_counter() - t0), file=sys.stderr, flush=True)
        return
    deck = decks.deck_3d_c1()
    ncells = decks.count_cells(deck)
    npatch = sum(len(l) for l in deck["levels"])
    print("bench: deck %s, %d patches, %d cells; building" % (deck["name"], npatch, ncells),
          file=sys.stderr, flush=True)
    h, u, f, s = build_problem(fac, decks, deck)
    if args.mode == "smoother":
        # a plain copy of known size calibrates the byte counters
        x = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        e = s.internal_vector("e")
        r = s.internal_vector("r")
        r.fill(1.0)
        s.smooth(e, r, 0, 20)
        s.smooth(e, r, len(deck["levels"]) - 1, 5)
        s.synchronize()
        print("bench: smoother mode done (copy 256 MiB, 20 coarsest sweeps, 5 finest sweeps)",
              file=sys.stderr, flush=True)
        return

    for w in range(args.warmup):
        u.fill(0.0)
        s.solve_system(u, f)
        print("bench: warmup %d: %d cycles, %.3f s" % (w, s.number_of_iterations(),
                                                        s.last_solve_seconds()),
              file=sys.stderr, flush=True)

    converged = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        u.fill(0.0)
        converged.append(s.solve_system(u, f))
    s.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    cycles = s.number_of_iterations()
    hist = s.history()
    fused = s.fused_coarsest_active()
    tile = fac.fused_tile_shapes(3)[s.fused_tile_variant()] if fused else None
    tuning = {"x".join(map(str, sh)): round(ms * 1e3, 2)
              for sh, ms in zip(fac.fused_tile_shapes(3), s.fused_tile_timings())}
    # one more solve with HIP events around every smoother launch (kept out
    # of the timed steps: the events add host work between launches)
    s.enable_kernel_timing(True)
    u.fill(0.0)
    s.solve_system(u, f)
    per_level = [s.kernel_timing(ln) for ln in range(len(deck["levels"]))]
    s.enable_kernel_timing(False)
    ms_per_step = elapsed * 1e3 / args.steps
    print("bench: %d steps, %.3f s/step, %d cycles, final residual %.3e" %
          (args.steps, ms_per_step / 1e3, cycles, hist[-1]), file=sys.stderr, flush=True)
    err = max_error(decks, deck, u)

    # smoother throughput (interior cells x full red+black sweeps / s) on
    # the coarsest level (the dominant operation) and on the finest level
    e = s.internal_vector("e")
    r = s.internal_vector("r")

    def smoother_rate(ln):
        n = sum(int(np.prod([hh - ll + 1 for ll, hh in zip(*b)])) for b in deck["levels"][ln])
        s.synchronize()
        t1 = time.perf_counter()
        s.smooth(e, r, ln, args.smoother_sweeps)
        s.synchronize()
        return n * args.smoother_sweeps / (time.perf_counter() - t1) / 1e9

    smoother_gcells = smoother_rate(0)
    finest_gcells = smoother_rate(len(deck["levels"]) - 1)

    # roofline of the smoother kernel that takes the most time: algorithmic
    # bytes of each launch (24 B per cell per full sweep at constant D: u
    # read, f read, u write) over its HIP-event duration
    levels = []
    for ln, (n, ms, b) in enumerate(per_level):
        levels.append(dict(level=ln, kernel=("k_box_rbgs (colour-fused sweep)" if ln == 0 and fused
                                             else "k_rbgs (one colour pass)"),
                           launches=n, total_ms=round(ms, 3),
                           avg_launch_us=round(ms * 1e3 / max(n, 1), 3),
                           bytes_per_launch=round(b / max(n, 1)),
                           achieved_GBs=round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else 0.0))
    dom = max(levels, key=lambda d: d["total_ms"])
    achieved = dom["achieved_GBs"]
    traffic = measured_traffic(tile) if dom["level"] == 0 and fused else None
    roofline = dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    traffic_source=("FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per launch "
                                    "(profiles/k_box_rbgs_c1_hbm_traffic.txt)"
                                    if traffic else None),
                    kernel="%s, level %d" % (dom["kernel"], dom["level"]),
                    launches=dom["launches"], avg_launch_us=dom["avg_launch_us"],
                    bytes_per_launch=dom["bytes_per_launch"], per_level=levels)

    # the same deck with the multigrid coarsest-level solver (the reference's
    # default coarse solver is hypre; "mg" is this library's stand-in for it)
    h2, u2, f2, s2 = build_problem(fac, decks, deck, MG_PARAMS)
    u2.fill(0.0)
    s2.solve_system(u2, f2)
    mg_times = []
    for k in range(args.steps):
        u2.fill(0.0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        ok2 = s2.solve_system(u2, f2)
        s2.synchronize()
        mg_times.append(time.perf_counter() - t2)
    mg = dict(s_per_solve=round(sum(mg_times) / len(mg_times), 4), cycles=s2.number_of_iterations(),
              converged=bool(ok2), final_residual=float(s2.history()[-1]),
              max_error_vs_exact=max_error(decks, deck, u2),
              coarse="mg V(2,2), rel tol 1e-8, <= 10 V-cycles",
              last_coarse_vcycles=s2.coarse_mg_stats()[0])
    print("bench: mg coarse solver: %.4f s/solve, %d cycles" % (mg["s_per_solve"], mg["cycles"]),
          file=sys.stderr, flush=True)

    # mixed precision: the coarsest-level solve of the error equation in FP32
    # (storage and arithmetic) inside the FP64 FAC iteration; errors per
    # level against the FP64 solution of the same solver (which equals the
    # oracle's to 1e-12 in the parity tests)
    def mixed(params, ref_u):
        hm, um, fm, sm = build_problem(fac, decks, deck, params)
        sm.set_coarse_precision(32)
        um.fill(0.0)
        sm.solve_system(um, fm)
        times = []
        for k in range(args.steps):
            um.fill(0.0)
            torch.cuda.synchronize()
            t3 = time.perf_counter()
            okm = sm.solve_system(um, fm)
            sm.synchronize()
            times.append(time.perf_counter() - t3)
        out = dict(s_per_solve=round(sum(times) / len(times), 4), cycles=sm.number_of_iterations(),
                   converged=bool(okm), final_residual=float(sm.history()[-1]),
                   error_vs_fp64_per_level=level_errors(deck, um, ref_u))
        return out
    mixed_c1 = dict(coarse="coarsest-level solve in FP32 storage and arithmetic; FAC "
                           "iteration, residuals and finer levels FP64",
                    redblack_coarse=mixed(PARAMS, u), mg_coarse=mixed(MG_PARAMS, u2))
    mixed_c1["redblack_coarse"]["fp64_s_per_solve"] = round(ms_per_step / 1e3, 4)
    mixed_c1["mg_coarse"]["fp64_s_per_solve"] = mg["s_per_solve"]
    print("bench: mixed precision C1 %s" % json.dumps(mixed_c1), file=sys.stderr, flush=True)

    c2 = None
    if not args.no_c2:
        print("bench: C2 deck (variable D), red-black and multigrid coarse solvers",
              file=sys.stderr, flush=True)
        deck2, rb2, _ = time_c2(fac, decks, torch, PARAMS)
        _, mg2, ref2 = time_c2(fac, decks, torch, MG_PARAMS)
        _, mg2m, mix2 = time_c2(fac, decks, torch, MG_PARAMS, bits=32)
        mg2m["error_vs_fp64_per_level"] = level_errors(deck2, mix2[1], ref2[1])
        mg2m["fp64_s_per_solve"] = mg2["s_per_solve"]
        c2 = dict(workload="c2-3d: 4 levels ratio 2, base 128^3, patch sides U{8..48} (seed 0), "
                           "D = exp(GRF, corr 0.1, sigma 1), f = 8 Gaussian sources (width 0.02)",
                  patches=sum(len(l) for l in deck2["levels"]), cells=decks.count_cells(deck2),
                  redblack_coarse=rb2, mg_coarse=mg2, mg_coarse_mixed_fp32=mg2m)
        print("bench: C2 %s" % json.dumps(c2), file=sys.stderr, flush=True)

    learned = None
    if not args.no_learned:
        learned = learned_section(torch, args.learned_seeds, args.learned_steps,
                                  args.learned_batch, log)
        log("bench: learned %s" % json.dumps(learned))

    cpu = None
    if not args.no_cpu_baseline:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            threads = min(16, os.cpu_count() or 1)
        print("bench: CPU baseline (%d cycles, %d threads)" % (args.cpu_cycles, threads),
              file=sys.stderr, flush=True)
        cpu = cpu_baseline(decks, deck, cycles, args.cpu_cycles, threads)
        mg["cpu_baseline"] = cpu_baseline(decks, deck, mg["cycles"], args.cpu_cycles, threads,
                                          mg=True)

    out = dict(metric="time to 1e-10 rel residual (s)", value=round(ms_per_step / 1e3, 4),
               unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=round(ms_per_step, 3), higher_is_better=False, scaling="strong",
               vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="%s: 3-D unit cube, 128^3 base, 3 levels ratio 2, "
                                    "|grad u| > 70th pct, boxes 16^3..64^3, D=1, sine RHS"
                                    % deck["name"],
                           patches=npatch, cells=ncells, levels=len(deck["levels"]),
                           solver="FAC V-cycle, RBGS 2+2, coarse redblack 500 sweeps tol "
                                  "1e-12, Ewing, conservative linear refine, rel tol 1e-10"),
               cycles=cycles, converged=all(converged), final_residual=float(hist[-1]),
               max_error_vs_exact=err,
               smoother_gcells_per_s=round(smoother_gcells, 2),
               finest_smoother_gcells_per_s=round(finest_gcells, 2),
               fused_coarsest=fused, fused_tile=tile, fused_tile_autotune_us_per_sweep=tuning,
               roofline=roofline, cpu_baseline=cpu, with_mg_coarse_solver=mg,
               mixed_precision=mixed_c1, c2=c2, learned_smoother=learned)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
