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


def train_pressure_preconditioner(problem, num_rhs=400, pairs_per_rhs=5, harvest_tol=1e-6, epochs=10, batch=16,
                                  lr=3e-3, seed=1, net_seed=2, layers=4, hidden=32, validate_max_iterations=600,
                                  device="cuda:0", log=None, base="sgs", muelu_params=None):
    """problem: dict with coords, conn, rho, p, u, gpdx (a lowmach_problem).
    base: the network's base preconditioner, "sgs" or "muelu" (the AMG V-cycle
    with muelu_params); the network is then meant for systems with that base.
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
    rec = {"train_mesh": "%d^3" % n_t, "pairs": len(pairs), "rhs": num_rhs, "pairs_per_rhs": pairs_per_rhs,
           "cg_tol": harvest_tol, "validation_fgmres_iters_per_epoch": net.validation, "kept_epoch": net.best_epoch,
           "harvest_s": t_harvest, "harvest_cg_iters_mean": float(np.mean(pairs.cg_iterations())),
           "train_s": t_train, "epochs": epochs, "batch": batch, "lr": lr, "epoch_loss": hist.tolist(),
           "layers": layers, "hidden": hidden, "weights": "bf16", "accumulate": "fp32", "base": base}
    lv.set_learned_preconditioner(None)
    del pairs, lv, lt
    return net, rec
