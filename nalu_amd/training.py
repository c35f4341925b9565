"""Training recipe of the learned pressure preconditioner (no reference
counterpart): harvest (residual, error) pairs from the plain FP64 CG + SGS
solver on a training mesh's pressure-Poisson system, train the message-passing
network with the library's own forward / backward / Adam kernels, and keep the
epoch whose network makes the mixed-precision FGMRES solve of the training
mesh's own pressure system converge in the fewest iterations.

Everything runs in libnalu_hip.so; this module only sequences the calls.
"""
import time

import numpy as np

from . import (HarvestedPairs, LearnedPreconditioner, LinearSystem, Mesh, NaluError, SolverConfig,
               assemble_continuity)


def _continuity_system(mesh, fields, cfg, dev, base="sgs", muelu_params=None):
    import torch
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=dev).contiguous()
    ls = LinearSystem(1, cfg)
    if muelu_params is not None:
        ls.set_muelu_params(muelu_params)
    if base != "sgs":
        ls.set_learned_base(base)
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    assemble_continuity(ls, mesh, t(fields["rho"]), t(fields["p"]), t(fields["u"]), t(fields["gpdx"]),
                        tau=fields.get("tau", 1.0), mode="gather")
    ls.loadComplete()
    return ls


def _self_pairs(net, pairs, mesh, problem, device, fraction, caps, seed, base, muelu_params):
    """Replace a fraction of the pairs by the residuals and errors of the
    network's own FGMRES iterates: seeded zero-mean Gaussian right-hand sides b,
    x* from FP64 CG + SGS to 1e-11, x_k after k FGMRES(50) iterations with the
    network (FP64) for each cap k, pair (b - A x_k, x* - x_k).  These are the
    residuals the network leaves behind in its own solves."""
    import torch
    lx = _continuity_system(mesh, problem, SolverConfig("cg", "sgs", tolerance=1e-11, max_iterations=5000), device)
    lfs = []
    for k in caps:
        lf = _continuity_system(mesh, problem, SolverConfig("fgmres", "learned", tolerance=1e-12, max_iterations=int(k),
                                                            kspace=50), device, base=base, muelu_params=muelu_params)
        lf.set_learned_preconditioner(net)
        lfs.append(lf)
    n = lx.sizes()[0]
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    nrep = int(round(fraction * len(pairs)))
    idx = torch.randperm(len(pairs), generator=g)[:nrep].tolist()
    q = 0
    y = torch.zeros(n, dtype=torch.float64, device=device)
    while q < nrep:
        b = torch.randn(n, generator=g, dtype=torch.float64).to(device)
        b -= b.mean()
        xs = torch.zeros_like(b)
        lx.solve_rhs(b, xs)
        for lf in lfs:
            if q >= nrep:
                break
            xk = torch.zeros_like(b)
            lf.solve_rhs(b, xk)
            lx.spmv(xk, y)
            pairs.set_pair(idx[q], (b - y).contiguous(), (xs - xk).contiguous())
            q += 1
    for lf in lfs:
        lf.set_learned_preconditioner(None)
    return nrep


def train_pressure_preconditioner(problem, num_rhs=400, pairs_per_rhs=5, harvest_tol=1e-6, epochs=10, batch=16,
                                  lr=3e-3, seed=1, net_seed=2, layers=4, hidden=32, validate_max_iterations=600,
                                  device="cuda:0", log=None, base="sgs", muelu_params=None, self_pair_rounds=0,
                                  self_pair_fraction=0.5, self_pair_caps=(3, 6, 10, 15, 25), self_pair_epochs=6,
                                  self_pair_lr=1e-3):
    """problem: dict with coords, conn, rho, p, u, gpdx (a lowmach_problem).
    base: the network's base preconditioner, "sgs" or "muelu" (the AMG V-cycle
    with muelu_params); the network is then meant for systems with that base.
    self_pair_rounds > 0: after the first training, that many rounds of
    replacing self_pair_fraction of the pairs by the network's own FGMRES
    residuals / errors (_self_pairs) and training self_pair_epochs more epochs
    at self_pair_lr, each round validated and selected like the first.
    Returns (network, record).  The record holds the harvest and training
    statistics and the validation FGMRES iterations of every epoch (10000 is
    added when the solve fell back to the base)."""
    import torch
    if num_rhs < 1 or pairs_per_rhs < 1 or epochs < 1:
        raise NaluError("train_pressure_preconditioner: num_rhs, pairs_per_rhs and epochs must be positive")
    say = log or (lambda msg: None)
    mesh = Mesh(problem["coords"], problem["conn"])
    n_t = round(problem["conn"].shape[0] ** (1.0 / 3.0))
    # the pairs come from CG + SGS histories; their stored feature is the base's output
    lt = _continuity_system(mesh, problem, SolverConfig("cg", "muelu" if base == "muelu" else "sgs"), device,
                            base=base, muelu_params=muelu_params)
    lt.precond_setup()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pairs = HarvestedPairs(lt, num_rhs=num_rhs, pairs_per_rhs=pairs_per_rhs, seed=seed, tol=harvest_tol)
    torch.cuda.synchronize()
    t_harvest = time.perf_counter() - t0
    say("harvested %d pairs at %d^3 in %.2f s" % (len(pairs), n_t, t_harvest))
    lv = _continuity_system(mesh, problem, SolverConfig("fgmres", "learned", tolerance=1e-8,
                                                        max_iterations=validate_max_iterations, kspace=50,
                                                        precision="mixed"), device, base=base,
                            muelu_params=muelu_params)
    lv.removeRhsMean()
    xv = torch.zeros(lv.sizes()[0], dtype=torch.float64, device=device)

    def validate(nn):
        lv.set_learned_preconditioner(nn)
        xv.zero_()
        lv.solve(xv)
        return lv.linearSolveIterations() + (10000 if lv.learned_fallback() else 0)

    net = LearnedPreconditioner(num_layers=layers, hidden=hidden, seed=net_seed)
    t0 = time.perf_counter()
    hist = net.train(lt, pairs, epochs=epochs, batch=batch, lr=lr, seed=seed, validate=validate)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    say("validation iterations per epoch %s, kept epoch %d" % (net.validation, net.best_epoch))
    say("trained %d epochs in %.2f s, losses %s" % (epochs, t_train, np.round(hist, 4).tolist()))
    first_validation, first_kept = list(net.validation), net.best_epoch  # the rounds below overwrite both
    rounds = []
    for rnd in range(int(self_pair_rounds)):
        t0 = time.perf_counter()
        nrep = _self_pairs(net, pairs, mesh, problem, device, self_pair_fraction, self_pair_caps, 1000 + rnd, base,
                           muelu_params)
        h2 = net.train(lt, pairs, epochs=self_pair_epochs, batch=batch, lr=self_pair_lr, seed=seed + 1 + rnd,
                       validate=validate)
        torch.cuda.synchronize()
        rounds.append({"pairs_replaced": nrep, "validation_fgmres_iters_per_epoch": net.validation,
                       "kept_epoch": net.best_epoch, "epoch_loss": h2.tolist(), "s": time.perf_counter() - t0})
        say("self-pair round %d: %d pairs replaced, validation %s, kept epoch %d" % (rnd + 1, nrep, net.validation,
                                                                                   net.best_epoch))
    rec = {"train_mesh": "%d^3" % n_t, "pairs": len(pairs), "rhs": num_rhs, "pairs_per_rhs": pairs_per_rhs,
           "cg_tol": harvest_tol, "validation_fgmres_iters_per_epoch": first_validation, "kept_epoch": first_kept,
           "harvest_s": t_harvest, "harvest_cg_iters_mean": float(np.mean(pairs.cg_iterations())),
           "train_s": t_train, "epochs": epochs, "batch": batch, "lr": lr, "epoch_loss": hist.tolist(),
           "layers": layers, "hidden": hidden, "weights": "bf16", "accumulate": "fp32", "base": base,
           "self_pair_rounds": rounds}
    lv.set_learned_preconditioner(None)
    del pairs, lv, lt
    return net, rec
