#!/usr/bin/env python3
"""FAC Poisson benchmark on one MI355X.

A step is one CellPoissonFACSolver::solveSystem from u = 0 to a relative
residual of 1e-10 (RMS residual <= 1e-10 * RMS f, FP64) on the
variable-coefficient deck: unit cube, 128^3 base, four levels at ratio 2
refined around eight Gaussian sources, patch sides drawn from U{8..48}
(2147 patches, 54 M cells), D = exp(g) with g a Gaussian random field
(correlation length 0.1, sigma 1), f = the eight Gaussians (width 0.02),
Dirichlet 0; red-black GS 2+2 sweeps, Ewing coarse-fine fluxes,
conservative linear prolongation, coarsest level by the library's geometric
multigrid (the red-black coarse solve, 500 sweeps, does not reach 1e-10 on
this deck and is reported beside it).

Prints one JSON line: time to solution (value, s), smoother throughput, the
roofline of the kernel with the largest share of the time (HIP events on the
solver's stream over one more, untimed, solve), the CPU baseline (the C
oracle on the host cores, all cores and one thread, a bounded sample of FAC
cycles projected to the GPU's cycle count) and, as extra keys, the 3-level
constant-coefficient deck, the reference's default coarse tolerance, mixed
precision and the learned smoother.
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

# BASELINE's solver: RBGS 2+2, relative residual 1e-10, Ewing, conservative
# linear prolongation.  Coarsest level: the library's multigrid (hypre's
# stopping test, relative 1e-8, at most 10 V-cycles) or red-black sweeps
# (up to 500 per coarse solve, pointwise tolerance 1e-12; the reference's
# default 1e-8 stalls the outer iteration, see the coarse_tolerance_1e-8 key).
PARAMS = dict(max_cycles=200, residual_tol=1e-300, relative_residual_tol=1e-10,
              num_pre_sweeps=2, num_post_sweeps=2, coarse_solver_choice="redblack",
              coarse_solver_tolerance=1e-12, coarse_solver_max_iterations=500,
              cf_discretization="Ewing", prolongation_method="CONSERVATIVE_LINEAR_REFINE")
MG_PARAMS = dict(PARAMS, coarse_solver_choice="mg", coarse_solver_tolerance=1e-8,
                 coarse_solver_max_iterations=10)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_description():
    """nproc and the lscpu fields that identify the host CPU."""
    out = dict(nproc=os.cpu_count())
    try:
        out["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    try:
        import subprocess
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket",
                     "Socket(s)", "NUMA node(s)", "CPU max MHz"):
                out[k] = v.strip()
    except Exception as e:  # lscpu missing: keep nproc
        out["lscpu"] = "unavailable (%s)" % e
    return out


def max_error(decks, deck, u):
    err = 0.0
    for ln, boxes in enumerate(deck["levels"]):
        for p, box in enumerate(boxes):
            shp = tuple(h - l + 3 for l, h in zip(*box))
            uu = u.get(ln, p).reshape(shp, order="F")[1:-1, 1:-1, 1:-1]
            err = max(err, float(np.max(np.abs(uu - decks.sine_exact(deck, ln, box)))))
    return err


def oracle_for(deck, params, f, D, threads):
    """The C oracle on the same hierarchy, keys and data (f, D copied from
    the GPU patch data)."""
    from oracle.oracle import OracleFAC
    o = OracleFAC(3, deck["levels"], deck["ratio"], deck["dlo"], deck["dhi"], [0.0] * 3,
                  [1.0] * 3)
    o.set_poisson(D is not None, 1.0, 0, 0.0)
    o.set_bc([1.0] * 6, [0.0] * 6, [0.0] * 6)
    o.set_params(max_cycles=params["max_cycles"], residual_tol=1e-300,
                 relative_residual_tol=1e-10, pre=2, post=2,
                 coarse_max_iterations=params["coarse_solver_max_iterations"],
                 coarse_tol=params["coarse_solver_tolerance"], ewing=True, nthreads=threads)
    o.set_coarse_mg(params["coarse_solver_choice"] == "mg")
    o.set_prolongation(params["prolongation_method"])
    for ln, boxes in enumerate(deck["levels"]):
        for p in range(len(boxes)):
            o.put("f", ln, p, f.get(ln, p))
            if D is not None:
                for a in range(3):
                    o.put("D", ln, p, D.get(ln, p, axis=a), axis=a)
    return o


def cpu_sample(deck, params, f, D, cycles_gpu, ncycles, threads):
    """ncycles FAC cycles of the oracle on `threads` host threads, projected
    to the GPU's cycle count (the initial composite residual is counted once,
    as in a whole solve)."""
    o = oracle_for(deck, dict(params, max_cycles=ncycles), f, D, threads)
    conv, it, res, hist = o.solve()
    secs = o.last_solve_seconds()
    o.close()
    per_cycle = secs / max(it, 1)
    return dict(value=round(per_cycle * cycles_gpu, 3), unit="s", cores=threads, kind="port",
                sample="%d FAC cycles of the same deck and keys with the C oracle on %d host "
                       "thread(s) (%.2f s, %.3f s/cycle incl. the initial residual), projected "
                       "to the %d cycles the GPU solve took" % (it, threads, secs, per_cycle,
                                                                cycles_gpu))


def measured_traffic(name):
    """HBM bytes per launch of a kernel from the committed counter summary
    (lines 'traffic <name> <bytes>'), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_per_launch.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic" and f[1] == name:
            return int(f[2])
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


def timed_solves(torch, s, u, f, steps):
    """Mean seconds of `steps` solves from u = 0 (after one warm-up solve
    that captures the cycle graph)."""
    u.fill(0.0)
    s.solve_system(u, f)
    times = []
    ok = False
    for _ in range(steps):
        u.fill(0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ok = s.solve_system(u, f)
        s.synchronize()
        times.append(time.perf_counter() - t0)
    return dict(s_per_solve=round(sum(times) / len(times), 4), cycles=s.number_of_iterations(),
                converged=bool(ok), final_residual=float(s.history()[-1]))


def kernel_roofline(s, u, f, deck, fused_box):
    """One more solve with HIP events around every smoother launch on the
    solver's stream; per level: launches, time and algorithmic bytes; the
    level whose smoother takes the most time gives the roofline."""
    s.enable_kernel_timing(True)
    u.fill(0.0)
    s.solve_system(u, f)
    per_level = [s.kernel_timing(ln) for ln in range(len(deck["levels"]))]
    s.enable_kernel_timing(False)
    levels = []
    for ln, (n, ms, b) in enumerate(per_level):
        if n == 0:
            continue
        kern = ("k_box_rbgs (colour-fused sweep)" if ln == 0 and fused_box
                else s.smoother_kernel_name(ln))
        levels.append(dict(level=ln, kernel=kern, launches=n, total_ms=round(ms, 3),
                           avg_launch_us=round(ms * 1e3 / max(n, 1), 3),
                           bytes_per_launch=round(b / max(n, 1)),
                           achieved_GBs=round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 else 0.0))
    dom = max(levels, key=lambda d: d["total_ms"])
    achieved = dom["achieved_GBs"]
    traffic = measured_traffic(dom["kernel"].split()[0] + "_L%d" % dom["level"])
    return dict(bound="hbm", achieved=achieved, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                traffic_source=("FETCH_SIZE + WRITE_SIZE per launch, scaled by a copy of known "
                                "size (profiles/hbm_traffic_per_launch.txt)" if traffic else None),
                kernel="%s, level %d" % (dom["kernel"], dom["level"]),
                launches=dom["launches"], avg_launch_us=dom["avg_launch_us"],
                bytes_per_launch=dom["bytes_per_launch"], per_level=levels)


def smoother_rate(s, deck, ln, sweeps):
    """Interior cells x full red+black sweeps per second (Gcell/s) of
    smoothErrorByRedBlack on level ln (the solver's error and residual)."""
    e = s.internal_vector("e")
    r = s.internal_vector("r")
    n = sum(int(np.prod([hh - ll + 1 for ll, hh in zip(*b)])) for b in deck["levels"][ln])
    s.smooth(e, r, ln, 2)
    s.synchronize()
    t1 = time.perf_counter()
    s.smooth(e, r, ln, sweeps)
    s.synchronize()
    return round(n * sweeps / (time.perf_counter() - t1) / 1e9, 2)


MFMA_BF16_PEAK_TFLOPS = 2500.0
# multiply-adds per cell of the network as specified (2 -> 16 -> 16 -> 1
# channels, 27 taps); the kernels pad every layer to 16 x 16 channels
LEARNED_MACS_PER_CELL = 27 * (2 * 16 + 16 * 16 + 16 * 1)
LEARNED_PADDED_MACS_PER_CELL = 27 * 16 * 16 * 3


def learned_section(torch, seeds, steps, batch, log=print, n=128, cycles=(1, 2, 3, 4, 5, 6, 7, 8)):
    """C4 (learned smoother): pairs generated on the device by the FAC solver
    on seeded variable-coefficient (D, f) draws (one-level n^3 hierarchies of
    16^3 patches, residuals after each cycle count in `cycles`, targets the
    FP64 block solutions; 50 draws at n = 128 and cycles 1..8 give 204,800),
    training on the GPU by the library's own forward/backward kernels, the
    matrix-core rate of the forward convolutions, the gain on unseen seeds, and
    whole solves with the correction on its own two-level problem and on C2."""
    from samrai_amd import decks, fac
    from samrai_amd import learned as SL
    t0 = time.perf_counter()
    train = []
    for sd in range(seeds):
        train.append(SL.fac_training_pairs(torch, sd, cycles=cycles, n=n, faces=False))
        if (sd + 1) % 10 == 0:
            log("bench: learned: pairs from %d of %d draws (%.1f s)" %
                (sd + 1, seeds, time.perf_counter() - t0))
    test = [SL.fac_training_pairs(torch, 1000 + s, cycles=cycles, n=n) for s in range(2)]
    cat = {k: torch.cat([p[k] for p in train]).contiguous() for k in ("r", "logd", "e")}
    del train[1:]
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

    def level_run(learned, sweeps):
        P.u.fill(0.0)
        hist = []
        t_it = 0.0
        for it in range(8):
            sv.synchronize()
            t2 = time.perf_counter()
            if learned:
                sv.residual(P.r, P.u, P.f, 0, False)
                SL.learned_correct(sv, net, P.u, P.r, P.D, 0)
            sv.smooth(P.u, P.f, 0, sweeps)
            sv.synchronize()
            t_it += time.perf_counter() - t2
            sv.residual(P.r, P.u, P.f, 0, False)
            hist.append(sv.rms_norm(P.r))
        return dict(rms_residual=[float("%.4e" % v) for v in hist],
                    ms_per_iteration=round(t_it / 8 * 1e3, 3))

    for learned in (False, True):
        level_run(learned, 2)  # warm-up
        level["learned_then_rbgs2" if learned else "rbgs2"] = level_run(learned, 2)
    # red-black alone at the learned iteration's cost (sweeps scaled by the
    # measured time per iteration)
    k_eq = max(2, int(round(2 * level["learned_then_rbgs2"]["ms_per_iteration"] /
                            level["rbgs2"]["ms_per_iteration"])))
    level["rbgs%d_equal_cost" % k_eq] = level_run(False, k_eq)
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
                                rollbacks=SL.learned_rollbacks(sv),
                                max_diff_vs_rbgs=float((uf - u_ref).abs().max()))
    SL.set_learned_smoother(sv, None)
    sv.close()
    P.close()
    log("bench: learned FAC solves %s" % json.dumps(fac_solves))
    # the same network on C2 (patches of 8..48 cells a side, covered by 16^3
    # windows): whole solves with red-black alone and with the correction
    deck2, h2, u2, f2, D2, s2 = decks.build_c2_problem(fac, torch, MG_PARAMS)
    nwin = max(sum(int(np.prod([-(-(hh - ll + 1) // 16) for ll, hh in zip(*b)])) for b in lv)
               for lv in deck2["levels"][1:])
    big = SL.LearnedSmoother(torch, nwin, seed=0)
    big.set_params(net.params())
    c2_solves = {}
    for name, om in (("rbgs", None), ("learned_omega_1", 1.0), ("learned_omega_0.5", 0.5)):
        SL.set_learned_smoother(s2, None if om is None else big, om or 1.0)
        u2.fill(0.0)
        s2.solve_system(u2, f2)  # graph capture
        u2.fill(0.0)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        ok = s2.solve_system(u2, f2)
        s2.synchronize()
        dt = time.perf_counter() - t2
        c2_solves[name] = dict(s_per_solve=round(dt, 4), cycles=s2.number_of_iterations(),
                               converged=bool(ok), final_residual=float(s2.history()[-1]),
                               learned_still_active=SL.learned_active(s2),
                               rollbacks=SL.learned_rollbacks(s2),
                               first_factors=[round(b / a, 4) for a, b in
                                              zip(s2.history()[:4], s2.history()[1:5])])
    SL.set_learned_smoother(s2, None)
    s2.close()
    big.close()
    log("bench: learned FAC solves on C2 %s" % json.dumps(c2_solves))
    out = dict(
        fac_solves=dict(workload="two levels, 64^3 + 128^3 in 16^3 patches (576 patches), "
                                 "D = exp(GRF), 8 Gaussian sources, unseen seed; RBGS 2+2, mg "
                                 "coarse, rel tol 1e-10", **fac_solves),
        fac_solves_c2=dict(workload="c2-3d-s0 (2147 patches of 8..48 cells a side), the "
                                    "same network over 16^3 windows (%d blocks on the largest "
                                    "level); RBGS 2+2, mg coarse, rel tol 1e-10" % nwin,
                           **c2_solves),
        level_smoothing=dict(workload="one-level 64^3 problem of 16^3 patches (unseen seed), "
                                      "8 iterations from u = 0", **level),
        workload="pairs from FAC iterates (cycles %s, 2 coarse RBGS sweeps per cycle) on "
                 "one-level %d^3 hierarchies of 16^3 patches, D = exp(GRF corr 0.1 sigma 1), "
                 "8 Gaussian sources, one (D, f) draw per seed; target = FP64 solution of the "
                 "block error equation (zero Dirichlet on the block faces, 400 SOR sweeps)"
                 % (",".join(str(c) for c in cycles), n),
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
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-cycles", type=int, default=2,
                    help="oracle cycles of the all-cores CPU sample (the one-thread sample "
                         "runs one cycle)")
    ap.add_argument("--no-cpu-1thread", action="store_true")
    ap.add_argument("--smoother-sweeps", type=int, default=40)
    ap.add_argument("--no-c1", action="store_true", help="skip the 3-level D = 1 deck (C1)")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-smoother (C4) section")
    ap.add_argument("--learned-seeds", type=int, default=50)
    ap.add_argument("--learned-steps", type=int, default=1500)
    ap.add_argument("--learned-batch", type=int, default=64)
    ap.add_argument("--learned-n", type=int, default=128,
                    help="cells per axis of each training draw's one-level hierarchy")
    ap.add_argument("--learned-cycles", default="1,2,3,4,5,6,7,8",
                    help="FAC cycle counts whose residuals become training pairs")
    ap.add_argument("--mode", choices=["solve", "smoother", "profile", "learned"],
                    default="solve",
                    help="smoother: a 256 MiB copy of known size and a few smoother calls on "
                         "--deck (for hardware-counter passes); profile: only the solves of "
                         "--deck with --coarse (for kernel traces); neither prints the "
                         "benchmark line")
    ap.add_argument("--deck", choices=["c1", "c2"], default="c2",
                    help="deck of --mode profile / smoother")
    ap.add_argument("--coarse", choices=["redblack", "mg"], default="mg",
                    help="coarsest-level solver of --mode profile")
    ap.add_argument("--mixed", type=int, choices=[64, 32, 16], default=64,
                    help="--mode profile: storage bits of the error vector on every level "
                         "(with 32 / 16 the coarsest-level solve runs in FP32)")
    ap.add_argument("--no-fused-levels", action="store_true",
                    help="--mode profile / smoother: patch-wise colour passes on the levels "
                         "above the coarsest instead of the colour-fused sweeps")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    from samrai_amd import decks, fac

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)

    learned_cycles = tuple(int(c) for c in args.learned_cycles.split(","))
    if args.mode == "learned":
        log("bench: learned %s" % json.dumps(learned_section(
            torch, args.learned_seeds, args.learned_steps, args.learned_batch, log,
            args.learned_n, learned_cycles)))
        return

    def problem(deck_name, params):
        if deck_name == "c2":
            deck, h, u, f, D, s = decks.build_c2_problem(fac, torch, params)
        else:
            deck = decks.deck_3d_c1()
            h, u, f, s = decks.build_sine_problem(fac, deck, params)
            D = None
        return deck, h, u, f, D, s

    if args.mode == "profile":
        deck, h, u, f, D, s = problem(args.deck, MG_PARAMS if args.coarse == "mg" else PARAMS)
        if args.no_fused_levels:
            s.set_fused_levels(False)
        if args.mixed != 64:
            s.set_coarse_precision(32)
            s.set_mixed_precision(args.mixed)
        # a plain copy of known size calibrates byte counters read from this mode
        x = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        torch.empty_like(x).copy_(x)
        del x
        torch.cuda.synchronize()
        for k in range(args.warmup + args.steps):
            u.fill(0.0)
            t0 = time.perf_counter()
            s.solve_system(u, f)
            s.synchronize()
            log("bench: profile %s %s solve %d: %d cycles, %.4f s" %
                (args.deck, args.coarse, k, s.number_of_iterations(), time.perf_counter() - t0))
        return
    if args.mode == "smoother":
        deck, h, u, f, D, s = problem(args.deck, MG_PARAMS)
        if args.no_fused_levels:
            s.set_fused_levels(False)
        # a plain copy of known size calibrates the byte counters
        x = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        e = s.internal_vector("e")
        r = s.internal_vector("r")
        r.fill(1.0)
        nl = len(deck["levels"])
        s.smooth(e, r, 0, 10)
        for ln in range(1, nl):
            s.smooth(e, r, ln, 2)
        # the composite residual of the error equation on every level, twice
        w = s.internal_vector("w")
        for _ in range(2):
            for ln in range(nl - 1, -1, -1):
                s.residual(w, e, r, ln, True)
        s.synchronize()
        log("bench: smoother mode done (copy 256 MiB, 10 coarsest sweeps, 2 sweeps per finer "
            "level, 2 composite residuals per level)")
        return

    # ---------------- the headline: C2, plain FP64, multigrid coarse solve
    deck, h, u, f, D, s = problem("c2", MG_PARAMS)
    ncells = decks.count_cells(deck)
    npatch = sum(len(l) for l in deck["levels"])
    log("bench: deck %s, %d patches, %d cells" % (deck["name"], npatch, ncells))
    for w in range(args.warmup):
        u.fill(0.0)
        s.solve_system(u, f)
        log("bench: warmup %d: %d cycles, %.3f s" % (w, s.number_of_iterations(),
                                                     s.last_solve_seconds()))
    converged = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        u.fill(0.0)
        converged.append(s.solve_system(u, f))
    s.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    cycles = s.number_of_iterations()
    hist = s.history()
    log("bench: %d steps, %.4f s/step, %d cycles, final residual %.3e" %
        (args.steps, ms_per_step / 1e3, cycles, hist[-1]))
    roofline = kernel_roofline(s, u, f, deck, s.fused_coarsest_active())
    log("bench: roofline %s" % json.dumps(roofline))
    nl = len(deck["levels"])
    gcells = {"level_%d" % ln: smoother_rate(s, deck, ln, args.smoother_sweeps)
              for ln in range(nl)}
    log("bench: smoother Gcell/s %s" % json.dumps(gcells))
    shapes = [None, (16, 8, 8), (8, 8, 8), (32, 4, 4), (16, 16, 4), (16, 8, 16)]
    level_smoother = {}
    for ln in range(1, nl):
        v = s.fused_level_variant(ln)
        level_smoother["level_%d" % ln] = ("colour passes" if v < 0 else
                                           "fused sweep, tile %dx%dx%d" % shapes[v + 1])
    log("bench: finer-level smoother (autotuned) %s" % json.dumps(level_smoother))
    u_ref = u  # the FP64 solution (mixed-precision errors are measured against it)

    # red-black coarse solve beside it (does not reach 1e-10 on this deck)
    _, _, ur, fr, Dr, sr = problem("c2", dict(PARAMS, max_cycles=60))
    rb = timed_solves(torch, sr, ur, fr, 1)
    rb["note"] = "max_cycles 60"
    log("bench: C2 red-black coarse solve %s" % json.dumps(rb))
    sr.close()
    del ur, fr, Dr

    # mixed precision inside the FP64 iteration: the coarsest-level solve in
    # FP32, then the error vector on every level in FP32 and in BF16 storage
    mixed = dict(fp64_s_per_solve=round(ms_per_step / 1e3, 4))
    runs = [("fp32_coarse_only", 64, 32,
             "coarsest-level solve (multigrid) in FP32 storage and arithmetic; FAC iteration, "
             "residuals, finer levels and u in FP64"),
            ("fp32_all_levels", 32, 32,
             "error vector in FP32 on every level: smoothing, its ghost fills and the "
             "prolongation in FP32 (D in an FP32 copy), coarsest-level solve in FP32; composite "
             "residuals, coarse-fine flux corrections, restriction, u += e and the "
             "convergence test in FP64"),
            ("bf16_all_levels", 16, 32,
             "error vector in BF16 storage on every level (FP32 arithmetic, D in a BF16 "
             "copy), coarsest-level solve in FP32; residuals, flux corrections, restriction, "
             "u += e and the convergence test in FP64")]
    for name, bits, coarse_bits, scope in runs:
        _, _, um, fm, Dm, sm = problem("c2", MG_PARAMS)
        sm.set_coarse_precision(coarse_bits)
        if bits != 64:
            sm.set_mixed_precision(bits)
        r = timed_solves(torch, sm, um, fm, max(1, min(args.steps, 5)))
        r["speedup_vs_fp64"] = round(ms_per_step / 1e3 / r["s_per_solve"], 3)
        r["error_vs_fp64_per_level"] = level_errors(deck, um, u_ref)
        r["scope"] = scope
        mixed[name] = r
        log("bench: C2 mixed precision %s %s" % (name, json.dumps(r)))
        sm.close()
        del um, fm, Dm

    c1 = None
    if not args.no_c1:
        deck1, _, u1, f1, _, s1 = problem("c1", PARAMS)
        c1 = dict(workload="c1-3d: 128^3 base, 3 levels ratio 2, |grad u| > 70th pct, Berger-"
                           "Rigoutsos boxes 16^3..64^3, D = 1, sine RHS",
                  patches=sum(len(l) for l in deck1["levels"]),
                  cells=decks.count_cells(deck1))
        c1["redblack_coarse"] = timed_solves(torch, s1, u1, f1, max(1, min(args.steps, 5)))
        c1["redblack_coarse"]["max_error_vs_exact"] = max_error(decks, deck1, u1)
        c1["redblack_coarse"]["roofline"] = kernel_roofline(s1, u1, f1, deck1,
                                                            s1.fused_coarsest_active())
        s1.close()
        _, _, u1m, f1m, _, s1m = problem("c1", MG_PARAMS)
        c1["mg_coarse"] = timed_solves(torch, s1m, u1m, f1m, max(1, min(args.steps, 5)))
        s1m.close()
        # the reference's default coarse tolerance (CellPoissonFACOps.C:594):
        # 1e-8 absolute max residual per coarse solve
        _, _, u1d, f1d, _, s1d = problem("c1", dict(PARAMS, coarse_solver_tolerance=1e-8,
                                                    max_cycles=100))
        c1["coarse_tolerance_1e-8"] = timed_solves(torch, s1d, u1d, f1d, 1)
        c1["coarse_tolerance_1e-8"]["history_every_10"] = [float("%.3e" % v)
                                                           for v in s1d.history()[::10]]
        s1d.close()
        log("bench: C1 %s" % json.dumps(c1))

    learned = None
    if not args.no_learned:
        learned = learned_section(torch, args.learned_seeds, args.learned_steps,
                                  args.learned_batch, log, args.learned_n, learned_cycles)
        log("bench: learned %s" % json.dumps(learned))

    cpu = None
    if not args.no_cpu_baseline:
        try:
            threads = min(16, len(os.sched_getaffinity(0)))
        except AttributeError:
            threads = min(16, os.cpu_count() or 1)
        log("bench: CPU baseline (%d cycles, %d threads)" % (args.cpu_cycles, threads))
        cpu = cpu_sample(deck, MG_PARAMS, f, D, cycles, args.cpu_cycles, threads)
        if not args.no_cpu_1thread:
            log("bench: CPU baseline (1 cycle, 1 thread)")
            cpu["one_thread"] = cpu_sample(deck, MG_PARAMS, f, D, cycles, 1, 1)
        cpu["host"] = host_description()
        log("bench: CPU %s" % json.dumps(cpu))

    out = dict(metric="time to 1e-10 rel residual (s)", value=round(ms_per_step / 1e3, 4),
               unit="s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=round(ms_per_step, 3), higher_is_better=False, scaling="strong",
               vs_baseline=None, dtype="f64", data="synthetic",
               config=dict(workload="%s: 3-D unit cube, 128^3 base, 4 levels ratio 2 around 8 "
                                    "Gaussian sources, patch sides U{8..48}, D = exp(GRF corr "
                                    "0.1 sigma 1), f = 8 Gaussians width 0.02" % deck["name"],
                           patches=npatch, cells=ncells, levels=nl,
                           patches_note="2147 patches for seed 0 (BASELINE: about 2500); "
                                        "the count follows from the seeded partition",
                           solver="FAC V-cycle, RBGS 2+2, Ewing, conservative linear refine, "
                                  "coarsest level multigrid V(2,2) rel tol 1e-8 <= 10 V-cycles, "
                                  "rel tol 1e-10"),
               cycles=cycles, converged=all(converged), final_residual=float(hist[-1]),
               smoother_gcells_per_s=gcells, finer_level_smoother=level_smoother,
               roofline=roofline, cpu_baseline=cpu,
               redblack_coarse=rb, mixed_precision=mixed, c1=c1, learned_smoother=learned)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
