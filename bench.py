#!/usr/bin/env python3
"""Benchmark: low-Mach Hex8 CVFEM assembly + pressure-Poisson solve on one MI355X.

Workload (default): the 128^3 low-Mach case -- unit cube, interior nodes
jittered +-0.1h, Taylor-Green velocity + N(0, 0.05^2) noise, density 1,
viscosity 1e-3, dt = gamma1 = 1.  One step is the hot path once:

  1. mdot at the 12 scs of every element (ComputeMdotElemAlgorithm)
  2. continuity system: zeroSystem, ContinuityAdvElemKernel assembly, loadComplete
     (by default 1 and 2 are one pass: the continuity row gather evaluates each
     scs's geometry and writes its mdot too; --separate-mdot runs them apart)
  3. momentum system (3 dof/node, 27-point block rows): zeroSystem, MomentumAdvDiffElemKernel
  4. pressure-Poisson solve: CG + multicolor SGS to 1e-8 FP64 true relative residual
     on the assembled continuity matrix with its zero-mean-projected rhs.

`value` is the assembly throughput (elements of steps 1-3 per second, Melem/s);
the pressure solve time, iterations and residual sit beside it.  Inputs are
resident in HBM before the timed region.  The dominant kernel's roofline and a
CPU baseline (the oracle's arithmetic built -O3 -march=x86-64-v3 -fopenmp, on every host core and on
one core, at the same size) are reported too.
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
METRIC = "Melem/s assembly + pressure solve s to 1e-8 resid, FP64 plain vs mixed+learned"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="elements per direction")
    ap.add_argument("--cpu-n", type=int, default=128, help="mesh size of the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 1 GiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--no-solver-compare", action="store_true",
                    help="skip the FP64 GMRES vs mixed + learned FGMRES pressure-solve comparison")
    ap.add_argument("--train-n", type=int, default=64, help="mesh of the learned preconditioner's training pairs")
    ap.add_argument("--train-epochs", type=int, default=10)
    ap.add_argument("--solve-reps", type=int, default=3)
    ap.add_argument("--muelu-degree", type=int, default=2, help="Chebyshev degree of the AMG smoother")
    ap.add_argument("--no-mixed", action="store_true", help="skip the FP32-assembly / mixed-precision block")
    ap.add_argument("--assembly-mode", default="gather", choices=["gather", "atomic", "tile", "etile"],
                    help="scatter strategy of the continuity and momentum assembly (default: row gather for continuity, row tiles for momentum)")
    ap.add_argument("--separate-mdot", action="store_true",
                    help="compute mdot in its own pass instead of inside the continuity row gather")
    ap.add_argument("--learned-apply-only", type=int, default=0, metavar="K",
                    help="only time K applications of an untrained learned preconditioner on the n^3 "
                         "continuity system (profiling aid), print one JSON line and exit")
    ap.add_argument("--amg-only", type=int, default=0, metavar="K",
                    help="only run K AMG set-ups + FP64 CG solves on the n^3 continuity system (lognormal(0, 0.3) "
                         "density; profiling aid), print one JSON line and exit")
    ap.add_argument("--stretched-only", action="store_true",
                    help="only the stretched-mesh pressure comparison (with --stretched-n), as one JSON line")
    ap.add_argument("--stretched-n", type=int, default=256,
                    help="run the stretched-mesh pressure comparison on an n^3 mesh (256: the stretched "
                         "configuration, about 2.5 min with the training; 0 skips it)")
    return ap.parse_args()


def algorithmic_bytes(N, E, P, rows, nnz1, nnz3):
    w, i = 8, 4
    return {
        "mdot": N * 11 * w + E * 32 + E * 12 * w,
        "continuity_assembly": N * 11 * w + E * 32 + P * w + N * w,
        "momentum_assembly": N * 7 * w + E * (32 + 12 * w) + 9 * P * w + 3 * N * w,
        "spmv": nnz1 * (w + i) + (rows + 1) * i + 2 * rows * w,
        # one SGS application = forward sweep over L + backward sweep over U: every
        # off-diagonal entry once, L/U row pointers, and per row r, 1/d (twice), z (3 accesses)
        "sgs_apply": (nnz1 - rows) * (w + i) + 2 * (rows + 1) * i + 6 * rows * w,
    }


# Algorithmic FP64 flops per unit of work (the same accounting as the byte
# counts: geometry ~5.4k flops per element, plus the kernel's own arithmetic;
# 2 flops per stored entry for a product, every off-diagonal entry once plus the
# diagonal scaling for one SGS application)
FP64_PEAK_TFLOPS = 78.6


def algorithmic_flops(E, rows, nnz1):
    return {
        "mdot": E * (5400 + 3000),
        "continuity_assembly": E * (5400 + 3000),
        "momentum_assembly": E * (5400 + 13000),
        "spmv": 2 * nnz1,
        "sgs_apply": 2 * (nnz1 - rows) + 4 * rows,
    }


def cpu_baseline(n_cpu):
    """The oracle's arithmetic compiled -O3 -march=x86-64-v3 -fopenmp (oracle/cpu_baseline.c),
    on the n_cpu^3 low-Mach case: assembly (zeroSystem + mdot + continuity +
    momentum, elements threaded, atomic sumInto) on every host core the process
    may use and on one core, and the CG + SGS pressure solve to 1e-8 on every
    core (SGS per thread block of rows, as Ifpack2 runs it per MPI rank)."""
    from nalu_amd.problems import lowmach_problem
    from oracle import cpu_baseline as CB
    P = lowmach_problem(n_cpu, seed=4321)
    E = P["conn"].shape[0]
    case = CB.LowMachCase(P)
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = CB.max_threads()
    CB.set_threads(threads)
    case.assemble()
    t_all = float(np.median([case.assemble() for _ in range(3)]))
    it, res, t_solve = case.pressure_solve(tol=1e-8)
    # one core on a bounded sample: the first 1/16 of the elements (same
    # per-element work), scaled to the whole mesh
    CB.set_threads(1)
    t_one = case.assemble(0, E // 16) * (E / (E // 16))
    CB.set_threads(threads)
    case.assemble()  # leave the full system assembled
    return {"value": E / t_all / 1e6, "unit": "Melem/s", "cores": threads, "kind": "port",
            "nproc": nproc, "one_core_melem_s": E / t_one / 1e6,
            "sample": "%d^3 low-Mach mesh (%d elements), same fields and seed as the GPU case: zeroSystem + mdot + "
                      "continuity + momentum assembly %.3f s on %d threads (median of 3), %.2f s on 1 thread (the "
                      "first 1/16 of the elements, scaled); CG + "
                      "per-thread-block SGS pressure solve to 1e-8: %d iterations, %.2f s, resid %.1e "
                      "(oracle arithmetic, gcc -O3 -march=x86-64-v3 -fopenmp)"
                      % (n_cpu, E, t_all, threads, t_one, it, t_solve, res),
            "assembly_s": t_all, "assembly_s_one_core": t_one, "pressure_solve_s": t_solve, "pressure_iters": it,
            "pressure_true_rel_resid": res}


def train_network(args, na, torch, devc, Pt, tag="", base="sgs", muelu_params=None, lr=3e-3):
    """Harvest 2000 (residual, error) pairs from plain FP64 CG + SGS on the
    pressure system of the training problem Pt (400 seeded Gaussian zero-mean
    right-hand sides x 5 pairs, CG to 1e-6: pairs are kept while the residual is
    above 1e-3 of its start, so the error is known to 1e-3), train the network
    and keep the epoch whose mixed FGMRES(50) solve of the training mesh's own
    pressure system (its assembled rhs, not a training pair) takes the fewest
    iterations (nalu_amd.training).  Returns (net, training record)."""
    from nalu_amd.training import train_pressure_preconditioner
    return train_pressure_preconditioner(
        Pt, num_rhs=400, pairs_per_rhs=5, harvest_tol=1e-6, epochs=args.train_epochs, batch=16, lr=lr, seed=1,
        net_seed=2, device=devc, log=lambda m: print("bench: %s%s" % (tag, m), file=sys.stderr, flush=True),
        base=base, muelu_params=muelu_params)


def solver_compare(args, na, torch, mesh, P, devc):
    """Pressure-Poisson solve to 1e-8 FP64 true relative residual on the same
    assembled continuity system: plain FP64 GMRES(50) + SGS against mixed
    precision FGMRES(50) with the learned preconditioner (trained on pairs
    harvested from FP64 CG + SGS on a train_n^3 mesh with lognormal(0, 0.3)
    density), plus the mixed FGMRES + SGS ablation and the smoothed-aggregation
    AMG preconditioner ("muelu": FP64 CG and mixed FGMRES; solve_s includes the
    hierarchy set-up, reported separately as amg.setup_s)."""
    from nalu_amd.problems import lowmach_problem
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=devc).contiguous()
    net, train_rec = train_network(args, na, torch, devc, lowmach_problem(args.train_n, rho_sigma=0.3, seed=4321))

    ev = lambda: torch.cuda.Event(enable_timing=True)
    res = {}
    sols = {}
    amg_prm = na.MueLuParams(cheb_degree=args.muelu_degree)
    for name, method, prec, precision in (("fp64_gmres_sgs", "gmres", "sgs", "fp64"),
                                          ("mixed_fgmres_learned", "fgmres", "learned", "mixed"),
                                          ("mixed_fgmres_sgs", "fgmres", "sgs", "mixed"),
                                          ("fp64_cg_muelu", "cg", "muelu", "fp64"),
                                          ("mixed_fgmres_muelu", "fgmres", "muelu", "mixed")):
        ls = na.LinearSystem(1, na.SolverConfig(method, prec, tolerance=1e-8, max_iterations=5000, kspace=50,
                                                precision=precision))
        if prec == "muelu":
            ls.set_muelu_params(amg_prm)
        ls.buildElemToNodeGraph(mesh)
        ls.finalizeLinearSystem()
        ls.set_stream(torch.cuda.current_stream().cuda_stream)
        ls.zeroSystem()
        na.assemble_continuity(ls, mesh, t(P["rho"]), t(P["p"]), t(P["u"]), t(P["gpdx"]), tau=P["tau"],
                               mode="gather")
        ls.loadComplete()
        ls.removeRhsMean()
        if prec == "learned":
            ls.set_learned_preconditioner(net)
        x = torch.zeros(ls.sizes()[0], dtype=torch.float64, device=devc)
        times = []
        for k in range(1 + args.solve_reps):
            x.zero_()
            ls.zeroSystem()
            na.assemble_continuity(ls, mesh, t(P["rho"]), t(P["p"]), t(P["u"]), t(P["gpdx"]), tau=P["tau"],
                                   mode="gather")
            ls.loadComplete()
            ls.removeRhsMean()
            torch.cuda.synchronize()
            a, b = ev(), ev()
            a.record()
            st = ls.solve(x)
            b.record()
            torch.cuda.synchronize()
            if k > 0:
                times.append(a.elapsed_time(b) * 1e-3)
        rel = ls.linearResidual() / max(ls.nonLinearResidual(), 1e-300)
        res[name] = {"solve_s": float(np.median(times)), "iters": ls.linearSolveIterations(),
                     "true_rel_resid": rel, "status": st}
        if prec == "learned":
            res[name]["fell_back_to_sgs"] = ls.learned_fallback()
        if prec == "muelu":
            res[name]["amg"] = dict(ls.amg_info(), cheb_degree=args.muelu_degree)
        sols[name] = x - x.mean()
        print("bench: %s %.4f s, %d iters, resid %.2e" % (name, res[name]["solve_s"], res[name]["iters"], rel),
              file=sys.stderr, flush=True)
    res["learned_apply"] = learned_apply_timing(args, na, torch, devc, net=net, mesh=mesh, P=P)
    ref = sols["fp64_gmres_sgs"]
    for name in ("mixed_fgmres_learned", "mixed_fgmres_sgs", "fp64_cg_muelu", "mixed_fgmres_muelu"):
        res[name]["pressure_rel_diff_vs_fp64"] = float(torch.linalg.norm(sols[name] - ref) / torch.linalg.norm(ref))
    res["speedup_fp64_cg_muelu_vs_fp64"] = res["fp64_gmres_sgs"]["solve_s"] / res["fp64_cg_muelu"]["solve_s"]
    res["speedup_mixed_muelu_vs_fp64"] = res["fp64_gmres_sgs"]["solve_s"] / res["mixed_fgmres_muelu"]["solve_s"]
    res["speedup_mixed_learned_vs_fp64"] = res["fp64_gmres_sgs"]["solve_s"] / res["mixed_fgmres_learned"]["solve_s"]
    res["speedup_mixed_sgs_vs_fp64"] = res["fp64_gmres_sgs"]["solve_s"] / res["mixed_fgmres_sgs"]["solve_s"]
    res["learned_training"] = train_rec
    return res, net


def timed(torch, fn, reps=1):
    """Median wall time of fn() with HIP events on the current stream."""
    ev = lambda: torch.cuda.Event(enable_timing=True)
    out = []
    for _ in range(reps):
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        out.append(a.elapsed_time(b) * 1e-3)
    return float(np.median(out))


def learned_apply_timing(args, na, torch, devc, net=None, mesh=None, P=None):
    """Time one application of the learned preconditioner (FP32, solver order)
    against one FP32 multicolor SGS application on the n^3 continuity system."""
    from nalu_amd.problems import lowmach_problem
    if mesh is None:
        P = lowmach_problem(args.n, seed=4321)
        mesh = na.Mesh(P["coords"], P["conn"])
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=devc).contiguous()
    ls = na.LinearSystem(1, na.SolverConfig("fgmres", "learned", precision="mixed"))
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.set_stream(torch.cuda.current_stream().cuda_stream)
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, t(P["rho"]), t(P["p"]), t(P["u"]), t(P["gpdx"]), tau=P["tau"], mode="gather")
    ls.loadComplete()
    ls.precond_setup()
    net = net or na.LearnedPreconditioner(num_layers=4, hidden=32, seed=2)
    nr = ls.sizes()[0]
    r = torch.randn(nr, dtype=torch.float32, device=devc)
    z = torch.zeros_like(r)
    reps = max(1, args.learned_apply_only or 20)
    net.apply(ls, r, z)
    t_net = timed(torch, lambda: [net.apply(ls, r, z) for _ in range(reps)]) / reps
    ls.spmv_f32(r, z)  # FP32 solver copy exists (mixed system)
    rd = r.double()
    zd = torch.zeros_like(rd)
    t_sgs = timed(torch, lambda: [ls.precond_apply_permuted(rd, zd) for _ in range(reps)]) / reps
    return {"learned_apply_ms": t_net * 1e3, "fp64_sgs_apply_ms": t_sgs * 1e3, "reps": reps,
            "note": "learned apply includes its FP32 SGS feature and the permutation in/out of nalu_gnn_apply"}


def mixed_block(args, na, torch, mesh, P, devc, stream):
    """FP32-arithmetic assembly (FP64 Jacobians) and mixed-precision solves on
    the same 128^3 low-Mach case, with the error of every field against the
    FP64 path (itself pinned to the FP64 oracle within 1e-12 by the tests)."""
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=devc).contiguous()
    rho, p, u, gp, mu = t(P["rho"]), t(P["p"]), t(P["u"]), t(P["gpdx"]), t(P["mu"])
    E = P["conn"].shape[0]
    res = {"errors_vs_fp64": {}, "solutions_vs_fp64": {}}
    md = {}
    sys_ = {}
    for prec in ("fp64", "mixed"):
        m = torch.zeros((E, 12), dtype=torch.float64, device=devc)
        na.compute_mdot(mesh, rho, p, u, gp, m, tau=P["tau"], stream=stream,
                        precision="fp32" if prec == "mixed" else "fp64")
        md[prec] = m
        cont = na.LinearSystem(1, na.SolverConfig("fgmres" if prec == "mixed" else "gmres", "sgs", tolerance=1e-8,
                                                  max_iterations=5000, precision=prec))
        mom = na.LinearSystem(3, na.SolverConfig("fgmres" if prec == "mixed" else "gmres", "sgs", tolerance=1e-8,
                                                 max_iterations=5000, precision=prec))
        for ls in (cont, mom):
            ls.buildElemToNodeGraph(mesh)
            ls.finalizeLinearSystem()
            ls.set_stream(stream)
        sys_[prec] = (cont, mom)

    def assemble(prec):
        cont, mom = sys_[prec]
        na.compute_mdot(mesh, rho, p, u, gp, md[prec], tau=P["tau"], stream=stream,
                        precision="fp32" if prec == "mixed" else "fp64")
        cont.zeroSystem()
        na.assemble_continuity(cont, mesh, rho, p, u, gp, tau=P["tau"], mode="gather")
        cont.loadComplete()
        mom.zeroSystem()
        na.assemble_momentum(mom, mesh, u, mu, md[prec], mode="tile")
        mom.loadComplete()

    def assemble_fused(prec):  # as the step does: mdot written by the continuity row gather
        cont, mom = sys_[prec]
        cont.zeroSystem()
        na.assemble_continuity(cont, mesh, rho, p, u, gp, tau=P["tau"], mode="gather", mdot_out=md[prec])
        cont.loadComplete()
        mom.zeroSystem()
        na.assemble_momentum(mom, mesh, u, mu, md[prec], mode="tile")
        mom.loadComplete()

    for prec in ("fp64", "mixed"):
        assemble_fused(prec)
        res["assembly_s_fused_mdot_" + prec] = timed(torch, lambda: assemble_fused(prec), reps=5)
        res["assembly_melem_s_fused_mdot_" + prec] = E / res["assembly_s_fused_mdot_" + prec] / 1e6
    for prec in ("fp64", "mixed"):
        assemble(prec)
        res["assembly_s_" + prec] = timed(torch, lambda: assemble(prec), reps=5)
    res["assembly_melem_s_mixed"] = E / res["assembly_s_mixed"] / 1e6
    res["assembly_melem_s_fp64"] = E / res["assembly_s_fp64"] / 1e6
    # per-field errors of the FP32 assembly
    err = res["errors_vs_fp64"]
    err["mdot_max_rel"] = float((md["mixed"] - md["fp64"]).abs().max() / md["fp64"].abs().max())
    for name, k in (("continuity", 0), ("momentum", 1)):
        a64, b64 = sys_["fp64"][k], sys_["mixed"][k]
        # host copies of both CSR systems (the graph is identical)
        _, _, vA, rA = a64.csr()
        _, _, vB, rB = b64.csr()
        err[name + "_matrix_max_rel"] = float(np.abs(vB - vA).max() / np.abs(vA).max())
        err[name + "_rhs_max_rel"] = float(np.abs(rB - rA).max() / np.abs(rA).max())
    # pressure solve (scalar, zero-mean); the coupled steady momentum system
    # without its time (mass) term is an advection-dominated operator that
    # GMRES(50) + SGS needs thousands of iterations for: its assembly is compared
    # here, the transient 3-rhs form is solved in momentum_3rhs()
    sol = {}
    for prec in ("fp64", "mixed"):
        cont, _ = sys_[prec]
        cont.removeRhsMean()
        xp = torch.zeros(cont.sizes()[0], dtype=torch.float64, device=devc)
        res["pressure_solve_s_" + prec] = timed(torch, lambda: cont.solve(xp))
        res["pressure_iters_" + prec] = cont.linearSolveIterations()
        res["pressure_true_rel_resid_" + prec] = cont.linearResidual() / max(cont.nonLinearResidual(), 1e-300)
        sol[prec] = xp - xp.mean()
    res["solutions_vs_fp64"]["pressure_rel_l2"] = float(torch.linalg.norm(sol["mixed"] - sol["fp64"]) /
                                                        torch.linalg.norm(sol["fp64"]))
    res["momentum_3rhs"] = momentum_3rhs(args, na, torch, mesh, P, devc, stream, md)
    return res


def momentum_3rhs(args, na, torch, mesh, P, devc, stream, md):
    """Momentum as three right-hand sides sharing one scalar matrix: one
    backward-Euler step (mass + advection-diffusion, dt = h so that the CFL
    number is about 1), the three velocity components solved with one matrix and
    one SGS set-up.  FP64 assembly + GMRES(50) against FP32-arithmetic assembly +
    mixed FGMRES(50), both to 1e-8; errors of the mixed fields against FP64."""
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=devc).contiguous()
    rho, u, gp, mu = t(P["rho"]), t(P["u"]), t(P["gpdx"]), t(P["mu"])
    N = P["coords"].shape[0]
    dt = 1.0 / args.n
    out = {"dt": dt, "rhs": 3, "errors_vs_fp64": {}}
    sols, csr = {}, {}
    for prec in ("fp64", "mixed"):
        ls = na.LinearSystem(1, na.SolverConfig("fgmres" if prec == "mixed" else "gmres", "sgs", tolerance=1e-8,
                                                max_iterations=2000, kspace=50, precision=prec))
        ls.buildElemToNodeGraph(mesh)
        ls.finalizeLinearSystem()
        ls.set_stream(stream)
        rhs3 = torch.zeros((3, N), dtype=torch.float64, device=devc)

        def asm():  # advection-diffusion and mass in one pass
            ls.zeroSystem()
            rhs3.zero_()
            na.assemble_momentum_with_mass_segregated(ls, mesh, u, mu, md[prec], (rho, rho, rho), (u, u, u), gp, dt,
                                                      rhs3)
            ls.loadComplete()

        asm()
        out["assembly_s_" + prec] = timed(torch, asm, reps=3)
        X3 = torch.zeros_like(rhs3)
        info = {}

        def solve():
            X3.zero_()
            info["r"] = ls.solve_multi(rhs3, X3)

        out["solve_s_" + prec] = timed(torch, solve, reps=3)
        st, its, rel = info["r"]
        out["iters_" + prec] = its
        out["true_rel_resid_" + prec] = rel
        sols[prec] = X3.clone()
        # the same three rhs as one stacked block-diagonal system (one matrix
        # read per SpMV / SGS sweep for the three components)
        XS = torch.zeros_like(rhs3)

        def solve_stacked():
            XS.zero_()
            info["s"] = ls.solve_stacked(rhs3, XS)

        out["stacked_solve_s_" + prec] = timed(torch, solve_stacked, reps=3)
        st_s, its_s, rel_s = info["s"]
        out["stacked_iters_" + prec] = its_s
        out["stacked_true_rel_resid_" + prec] = rel_s
        out["stacked_vs_separate_rel_l2_" + prec] = float(torch.linalg.norm(XS - X3) / torch.linalg.norm(X3))
        csr[prec] = (ls.csr(), rhs3.cpu().numpy())
    err = out["errors_vs_fp64"]
    (_, _, vA, _), bA = csr["fp64"]
    (_, _, vB, _), bB = csr["mixed"]
    err["matrix_max_rel"] = float(np.abs(vB - vA).max() / np.abs(vA).max())
    for i, c in enumerate("uvw"):
        err["rhs_%s_max_rel" % c] = float(np.abs(bB[i] - bA[i]).max() / np.abs(bA[i]).max())
        err["delta_%s_rel_l2" % c] = float(torch.linalg.norm(sols["mixed"][i] - sols["fp64"][i]) /
                                           torch.linalg.norm(sols["fp64"][i]))
    # the reference's form of the same step: one numDof-3 system whose lhs carries
    # the transpose-gradient stress coupling, FP64 GMRES(50) + SGS to 1e-8
    lc = na.LinearSystem(3, na.SolverConfig("gmres", "sgs", tolerance=1e-8, max_iterations=2000, kspace=50))
    lc.buildElemToNodeGraph(mesh)
    lc.finalizeLinearSystem()
    lc.set_stream(stream)

    def asm_coupled_two_passes():
        lc.zeroSystem()
        na.assemble_momentum(lc, mesh, u, mu, md["fp64"])
        na.assemble_momentum_mass(lc, mesh, (rho, rho, rho), (u, u, u), gp, dt)
        lc.loadComplete()

    def asm_coupled():  # advection-diffusion and mass in one element pass
        lc.zeroSystem()
        na.assemble_momentum_with_mass(lc, mesh, u, mu, md["fp64"], (rho, rho, rho), (u, u, u), gp, dt)
        lc.loadComplete()

    asm_coupled_two_passes()
    out["coupled_assembly_two_passes_s_fp64"] = timed(torch, asm_coupled_two_passes, reps=3)
    asm_coupled()
    out["coupled_assembly_s_fp64"] = timed(torch, asm_coupled, reps=3)
    xc = torch.zeros(3 * N, dtype=torch.float64, device=devc)

    def solve_coupled():
        xc.zero_()
        lc.solve(xc)

    out["coupled_solve_s_fp64"] = timed(torch, solve_coupled, reps=3)
    out["coupled_iters_fp64"] = lc.linearSolveIterations()
    out["coupled_true_rel_resid_fp64"] = lc.linearResidual() / max(lc.nonLinearResidual(), 1e-300)
    # the segregated form lags the coupling in its lhs: the increments differ at this
    # order while both systems share the rhs (the full residual)
    out["coupled_vs_segregated_delta_rel_l2"] = float(torch.linalg.norm(xc.reshape(N, 3).T - sols["fp64"]) /
                                                      torch.linalg.norm(sols["fp64"]))
    print("bench: momentum coupled numDof 3: assembly %.4f s (two passes %.4f s), fp64 solve %.4f s %d it., "
          "resid %.2e" % (out["coupled_assembly_s_fp64"], out["coupled_assembly_two_passes_s_fp64"],
                          out["coupled_solve_s_fp64"], out["coupled_iters_fp64"], out["coupled_true_rel_resid_fp64"]),
          file=sys.stderr, flush=True)
    del lc
    out["speedup_mixed_vs_fp64"] = out["solve_s_fp64"] / out["solve_s_mixed"]
    out["speedup_stacked_vs_separate_fp64"] = out["solve_s_fp64"] / out["stacked_solve_s_fp64"]
    out["speedup_stacked_mixed_vs_separate_fp64"] = out["solve_s_fp64"] / out["stacked_solve_s_mixed"]
    print("bench: momentum 3 rhs stacked: fp64 %.4f s %d it., mixed %.4f s %d it." %
          (out["stacked_solve_s_fp64"], out["stacked_iters_fp64"], out["stacked_solve_s_mixed"],
           out["stacked_iters_mixed"]), file=sys.stderr, flush=True)
    print("bench: momentum 3 rhs: fp64 %.4f s %s, mixed %.4f s %s" % (out["solve_s_fp64"], out["iters_fp64"],
                                                                     out["solve_s_mixed"], out["iters_mixed"]),
          file=sys.stderr, flush=True)
    return out


def amg_timing(args, na, torch, devc):
    """K rounds of: assemble the n^3 continuity system, AMG set-up, FP64 CG + AMG
    solve to 1e-8 (set-up and solve timed apart)."""
    from nalu_amd.problems import lowmach_problem
    P = lowmach_problem(args.n, rho_sigma=0.3, seed=4321)
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=devc).contiguous()
    mesh = na.Mesh(P["coords"], P["conn"])
    ls = na.LinearSystem(1, na.SolverConfig("cg", "muelu", tolerance=1e-8, max_iterations=500))
    ls.set_muelu_params(na.MueLuParams(cheb_degree=args.muelu_degree))
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.set_stream(torch.cuda.current_stream().cuda_stream)
    x = torch.zeros(ls.sizes()[0], dtype=torch.float64, device=devc)
    setup, total = [], []
    for _ in range(args.amg_only):
        x.zero_()
        ls.zeroSystem()
        na.assemble_continuity(ls, mesh, t(P["rho"]), t(P["p"]), t(P["u"]), t(P["gpdx"]), tau=1.0, mode="gather")
        ls.loadComplete()
        ls.removeRhsMean()
        total.append(timed(torch, lambda: ls.solve(x)))
        setup.append(ls.amg_info()["setup_s"])
    return {"amg_setup_s": setup, "cg_amg_solve_s_incl_setup": total, "iters": ls.linearSolveIterations(),
            "hierarchy": ls.amg_info()}


def stretched_compare(args, na, torch, devc):
    """Stretched-mesh pressure-Poisson: geometric ratio 1.05 from all faces
    (aspect ratio capped at 100), jitter 0.05h, density lognormal(0, 0.3).
    Plain FP64 GMRES(50) + SGS against mixed FGMRES(50) + learned preconditioner
    to the same 1e-8 FP64 relative residual (one timed solve each).  On these
    cells SGS barely converges, so the network is trained on the AMG V-cycle as
    its base (its input feature, initial output and fallback); the AMG alone
    (mixed FGMRES and FP64 CG) and mixed FGMRES + SGS are the ablations."""
    from nalu_amd.problems import lowmach_problem
    n = args.stretched_n
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=devc).contiguous()
    mp = na.MueLuParams(cheb_degree=args.muelu_degree)
    # the network for stretched cells is trained on a stretched train_n^3 mesh whose
    # per-cell ratio reaches the same growth over the same fraction of the box
    ratio_t = 1.05 ** (n / args.train_n)
    net, rec = train_network(args, na, torch, devc,
                             lowmach_problem(args.train_n, jitter=0.05, rho_sigma=0.3, seed=778, stretch=ratio_t,
                                             max_aspect=100.0), tag="stretched ", base="muelu", muelu_params=mp,
                             lr=5e-4)
    rec["stretch_ratio"] = ratio_t
    P = lowmach_problem(n, jitter=0.05, rho_sigma=0.3, seed=777, stretch=1.05, max_aspect=100.0)
    hx = np.diff(P["coords"][:n + 1, 0])  # the boundary line j = k = 0 carries the unjittered axis positions
    mesh = na.Mesh(P["coords"], P["conn"])
    out = {"mesh": "%d^3" % n, "elements": int(P["conn"].shape[0]), "stretch_ratio": 1.05,
           "max_to_min_spacing": float(hx.max() / hx.min()), "learned_training": rec}
    for name, method, prec, precision in (("fp64_gmres_sgs", "gmres", "sgs", "fp64"),
                                          ("mixed_fgmres_learned", "fgmres", "learned", "mixed"),
                                          ("mixed_fgmres_muelu", "fgmres", "muelu", "mixed"),
                                          ("fp64_cg_muelu", "cg", "muelu", "fp64"),
                                          ("mixed_fgmres_sgs", "fgmres", "sgs", "mixed")):
        ls = na.LinearSystem(1, na.SolverConfig(method, prec, tolerance=1e-8, max_iterations=20000, kspace=50,
                                                precision=precision))
        if prec in ("muelu", "learned"):
            ls.set_muelu_params(mp)
        if prec == "learned":
            ls.set_learned_base("muelu")
        ls.buildElemToNodeGraph(mesh)
        ls.finalizeLinearSystem()
        ls.set_stream(torch.cuda.current_stream().cuda_stream)
        ls.zeroSystem()
        na.assemble_continuity(ls, mesh, t(P["rho"]), t(P["p"]), t(P["u"]), t(P["gpdx"]), tau=1.0, mode="gather")
        ls.loadComplete()
        ls.removeRhsMean()
        if prec == "learned":
            ls.set_learned_preconditioner(net)
        x = torch.zeros(ls.sizes()[0], dtype=torch.float64, device=devc)
        if prec in ("muelu", "learned"):  # warm the set-up path once (first-touch allocations)
            ls.solve(x)
            x.zero_()
        dt = timed(torch, lambda: ls.solve(x))
        rel = ls.linearResidual() / max(ls.nonLinearResidual(), 1e-300)
        out[name] = {"solve_s": dt, "iters": ls.linearSolveIterations(), "true_rel_resid": rel}
        if prec == "learned":
            out[name]["fell_back_to_base"] = ls.learned_fallback()
            out[name]["base"] = "muelu"
        if prec in ("muelu", "learned"):
            out[name]["amg"] = ls.amg_info()
        print("bench: stretched %s %.3f s, %d iters, resid %.2e" % (name, dt, out[name]["iters"], rel),
              file=sys.stderr, flush=True)
        del ls, x
        torch.cuda.empty_cache()
    base = out["fp64_gmres_sgs"]["solve_s"]
    for k in ("mixed_fgmres_learned", "mixed_fgmres_muelu", "fp64_cg_muelu", "mixed_fgmres_sgs"):
        out["speedup_%s_vs_fp64_gmres_sgs" % k] = base / out[k]["solve_s"]
    out["learned_vs_its_amg_base"] = out["mixed_fgmres_muelu"]["solve_s"] / out["mixed_fgmres_learned"]["solve_s"]
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (the path would shard by domain "
              "decomposition: element slabs per GPU, owned / shared rows, one halo exchange per SpMV).",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import nalu_amd as na
    from nalu_amd.problems import lowmach_problem

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    devc = torch.device("cuda:0")
    n = args.n
    if args.learned_apply_only > 0:
        print(json.dumps(learned_apply_timing(args, na, torch, devc)))
        return
    if args.stretched_only:
        print(json.dumps(stretched_compare(args, na, torch, devc)))
        return
    if args.amg_only > 0:
        print(json.dumps(amg_timing(args, na, torch, devc)))
        return
    print("bench: generating %d^3 low-Mach problem" % n, file=sys.stderr, flush=True)
    P = lowmach_problem(n, seed=4321)
    N, E = P["coords"].shape[0], P["conn"].shape[0]
    mesh = na.Mesh(P["coords"], P["conn"])
    cont = na.LinearSystem(1, na.SolverConfig("cg", "sgs", tolerance=1e-8, max_iterations=5000))
    mom = na.LinearSystem(3, na.SolverConfig("gmres", "sgs"))
    for ls in (cont, mom):
        ls.buildElemToNodeGraph(mesh)
        ls.finalizeLinearSystem()
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=devc).contiguous()
    rho, p, u, gp, mu = t(P["rho"]), t(P["p"]), t(P["u"]), t(P["gpdx"]), t(P["mu"])
    mdot = torch.zeros((E, 12), dtype=torch.float64, device=devc)
    # Gpdx: projected nodal gradient of p (interior Green-Gauss + exposed-face
    # closure) on the dual nodal volume, computed on the GPU every step
    dual_vol = torch.zeros(N, dtype=torch.float64, device=devc)
    na.compute_dual_volume(mesh, dual_vol)
    gp = torch.zeros((N, 3), dtype=torch.float64, device=devc)
    rows, nnz1 = cont.sizes()
    _, nnz3 = mom.sizes()
    Pn = nnz1  # scalar node-pair nonzeros
    x = torch.zeros(rows, dtype=torch.float64, device=devc)
    b = torch.randn(rows, dtype=torch.float64, device=devc)
    # the library runs on torch's current stream, so torch events bracket its kernels
    stream = torch.cuda.current_stream().cuda_stream
    for ls in (cont, mom):
        ls.set_stream(stream)

    if args.copy_probe:
        src = torch.ones(1 << 27, dtype=torch.float64, device=devc)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ev = lambda: torch.cuda.Event(enable_timing=True)
    cmode = args.assembly_mode
    mmode = "tile" if args.assembly_mode == "gather" else args.assembly_mode  # no row gather for numDof 3
    fused = cmode == "gather" and not args.separate_mdot

    def step(timers=None):
        e = [ev() for _ in range(6)] if timers is not None else None
        if e: e[5].record()
        na.compute_nodal_gradient(mesh, p, dual_vol, gp, stream=stream)
        if e: e[0].record()
        if not fused:
            na.compute_mdot(mesh, rho, p, u, gp, mdot, tau=P["tau"], stream=stream)
        if e: e[1].record()
        cont.zeroSystem()
        if fused:  # continuity + mdot from one evaluation of the scs geometry
            na.assemble_continuity(cont, mesh, rho, p, u, gp, tau=P["tau"], mode="gather", mdot_out=mdot)
        else:
            na.assemble_continuity(cont, mesh, rho, p, u, gp, tau=P["tau"], mode=cmode)
        cont.loadComplete()
        if e: e[2].record()
        mom.zeroSystem()
        na.assemble_momentum(mom, mesh, u, mu, mdot, mode=mmode)
        mom.loadComplete()
        if e: e[3].record()
        # zero-mean projection of the continuity rhs (singular Neumann pressure system)
        cont.removeRhsMean()
        st = cont.solve(x)
        it = cont.linearSolveIterations()
        rel = cont.linearResidual() / max(cont.nonLinearResidual(), 1e-300)
        if e: e[4].record()
        if timers is not None:
            timers.append((e, it, rel, st))
        return it, rel, st

    print("bench: warmup %d" % args.warmup, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    gp_ref = t(P["gpdx"])
    gp_rel = float(torch.linalg.norm(gp - gp_ref) / torch.linalg.norm(gp_ref))
    timers = []
    print("bench: timing %d steps" % args.steps, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(timers)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    ph = {"nodal_gradient": [], "mdot": [], "continuity_assembly": [], "momentum_assembly": [], "pressure_solve": []}
    for (e, it, rel, st) in timers:
        ph["nodal_gradient"].append(e[5].elapsed_time(e[0]) * 1e-3)
        ph["mdot"].append(e[0].elapsed_time(e[1]) * 1e-3)
        ph["continuity_assembly"].append(e[1].elapsed_time(e[2]) * 1e-3)
        ph["momentum_assembly"].append(e[2].elapsed_time(e[3]) * 1e-3)
        ph["pressure_solve"].append(e[3].elapsed_time(e[4]) * 1e-3)
    med = {k: float(np.median(v)) for k, v in ph.items()}
    iters = timers[-1][1]
    resid = timers[-1][2]
    t_asm = med["mdot"] + med["continuity_assembly"] + med["momentum_assembly"]

    # kernel-level timing of the solver building blocks, HIP events on the stream they run on
    def time_op(fn, reps):
        fn()
        torch.cuda.synchronize()
        a, bb = ev(), ev()
        a.record()
        for _ in range(reps):
            fn()
        bb.record()
        torch.cuda.synchronize()
        return a.elapsed_time(bb) * 1e-3 / reps

    y = torch.zeros_like(x)
    # the solver's own kernels, on its color-ordered copy of the matrix
    t_spmv = time_op(lambda: cont.spmv_permuted(b, y), 20)
    t_sgs = time_op(lambda: cont.precond_apply_permuted(b, y), 10)
    cp, _ = cont.coloring()
    ncol = len(cp) - 1

    byt = algorithmic_bytes(N, E, Pn, rows, nnz1, nnz3)
    flp = algorithmic_flops(E, rows, nnz1)
    if fused:  # the continuity pass also writes mdot [E][12]
        byt["continuity_assembly"] += 12 * E * 8
    per_step = {
        "mdot": (med["mdot"], byt["mdot"], 1),
        "continuity_assembly": (med["continuity_assembly"], byt["continuity_assembly"], 1),
        "momentum_assembly": (med["momentum_assembly"], byt["momentum_assembly"], 1),
        "sgs_apply": (t_sgs * (iters + 1), byt["sgs_apply"], 2 * ncol - 1),
        "spmv": (t_spmv * (iters + 1), byt["spmv"], 1),
    }
    kname = {"gather": ("k_scalar_gather<ContinuityOp>", "k_momentum_tile"),
             "atomic": ("k_scalar_atomic<ContinuityOp>", "k_momentum_atomic"),
             "tile": ("k_scalar_tile<ContinuityOp>", "k_momentum_tile"),
             "etile": ("k_scalar_items<ContinuityEt> + k_scalar_etile", "k_momentum_items + k_momentum_etile")}
    if fused:
        del per_step["mdot"]
    kern_name = {"mdot": "k_mdot_elem", "continuity_assembly": kname[cmode][0] + (" (+ mdot)" if fused else ""),
                 "momentum_assembly": kname[mmode][1], "sgs_apply": "k_sell_fwd + k_sell_bwd",
                 "spmv": "k_sell_spmv"}
    def unit_t(k):
        return {"sgs_apply": t_sgs, "spmv": t_spmv}.get(k, per_step[k][0])

    dom = max(per_step, key=lambda k: per_step[k][0])
    unit_time = {"mdot": med["mdot"], "continuity_assembly": med["continuity_assembly"],
                 "momentum_assembly": med["momentum_assembly"], "sgs_apply": t_sgs, "spmv": t_spmv}[dom]
    launches = per_step[dom][2]
    achieved = byt[dom] / unit_time / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s_%d.json" % (dom, n))
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    compare = net = None
    if not args.no_solver_compare:
        compare, net = solver_compare(args, na, torch, mesh, P, devc)

    mixed = None
    if not args.no_mixed:
        print("bench: mixed-precision assembly and solves", file=sys.stderr, flush=True)
        mixed = mixed_block(args, na, torch, mesh, P, devc, stream)

    stretched = None
    if args.stretched_n > 0:
        del cont, mom
        torch.cuda.empty_cache()
        print("bench: stretched %d^3 pressure comparison" % args.stretched_n, file=sys.stderr, flush=True)
        stretched = stretched_compare(args, na, torch, devc)

    cpu = None
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (threaded oracle arithmetic, %d^3)" % args.cpu_n, file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_n)

    out = {
        "metric": METRIC,
        "value": E / t_asm / 1e6,
        "unit": "Melem/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded LCG mesh jitter and velocity noise, analytic Taylor-Green fields)",
        "config": {"workload": "low-Mach momentum + continuity assembly and CG+SGS pressure-Poisson solve, "
                               "%d^3 Hex8, jitter 0.1h, plain FP64" % n, "assembly_mode": cmode,
                   "mdot": "written by the continuity row gather" if fused else "separate pass",
                   "elements": E, "nodes": N, "continuity_nnz": nnz1, "momentum_nnz": nnz3,
                   "parallelism": "single GPU"},
        "assembly_s": t_asm,
        "phase_s": med,
        "pressure_solve_s": med["pressure_solve"],
        "pressure_iters": iters,
        "pressure_true_rel_resid": resid,
        "sgs_colors": ncol,
        "gpdx": {"source": "projected nodal gradient of p on the GPU each step (interior + exposed-face closure)",
                 "rel_l2_vs_analytic_gradient": gp_rel},
        "pressure_solve_compare": compare,
        "mixed_precision": mixed,
        "stretched_pressure_compare": stretched,
        "kernels": {
            k: {"time_s": per_step[k][0], "unit_time_s": unit_t(k),
                "algorithmic_bytes": byt[k],
                "GBps": byt[k] / unit_t(k) / 1e9,
                "hbm_frac": byt[k] / unit_t(k) / 1e9 / HBM_PEAK_GBS,
                "algorithmic_fp64_flops": flp[k],
                "fp64_TFLOPs": flp[k] / unit_t(k) / 1e12,
                "fp64_frac": flp[k] / unit_t(k) / 1e12 / FP64_PEAK_TFLOPS}
            for k in per_step},
        "roofline": {"bound": "hbm", "kernel": kern_name[dom], "unit_of_work": dom,
                     "launches_per_unit": launches,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
