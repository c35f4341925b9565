"""GPU tests of the learned preconditioner (HIP kernels against the CPU oracle).

Bars: the SGS input feature matches the oracle's multicolor SGS to FP32
accuracy (1e-5 relative); the network output, the loss and the gradient match
the oracle with the kernels' BF16 rounding points emulated (output 5e-3,
loss 2e-3 relative; each parameter block's gradient 2e-2 relative -- the
kernels accumulate in FP32, so a hidden value can round to the neighbouring
BF16 number, a 2^-8 relative step, where the oracle's FP64 sum does not); the
untrained network reproduces SGS; harvested pairs satisfy A e = r to the
1e-3 the harvest guarantees; training
lowers the loss and cuts FGMRES iterations against the same solver with SGS
alone; a network with non-finite weights falls back to SGS and still converges.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nalu_amd as na  # noqa: E402
from nalu_amd.problems import lowmach_problem, poisson_rhs  # noqa: E402
from oracle import gnn as G  # noqa: E402

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def poisson_system(n, cfg=None, seed=4321):
    P = lowmach_problem(n, rho_sigma=0.3, seed=seed)
    mesh = na.Mesh(P["coords"], P["conn"])
    ls = na.LinearSystem(1, cfg or na.SolverConfig("fgmres", "learned", tolerance=1e-8, max_iterations=2000,
                                                   kspace=30, precision="mixed"))
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), tau=1.0,
                           mode="gather")
    ls.loadComplete()
    ls.removeRhsMean()
    ls.precond_setup()
    return P, mesh, ls


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_forward_loss_grad_match_oracle(oracle):
    print("case: learned loss/grad vs oracle, n=6", flush=True)
    _, _, ls = poisson_system(6)
    rp, ci, v, _ = ls.csr()
    n = len(rp) - 1
    gr = G.Graph(rp, ci, v)
    net = na.LearnedPreconditioner(num_layers=4, seed=3)
    o = G.layout(4)
    rng = np.random.default_rng(5)
    th = net.get_params().astype(np.float64)
    th[o["w_out"]:o["c"]] = rng.normal(0, 0.3, 32)
    th[o["c"]] = 0.3
    for l in range(4):
        th[o[f"b{l}"]:o[f"b{l}"] + 32] = rng.normal(0, 0.1, 32)
    th = th.astype(np.float32)
    net.set_params(th)
    B = 3
    e = rng.standard_normal((B, n)).astype(np.float32)
    r = np.stack([oracle.spmv(rp, ci, v, e[b].astype(np.float64)) for b in range(B)]).astype(np.float32)
    zs_d = torch.zeros(B, n, dtype=torch.float32, device=DEV)
    out_d = torch.zeros(B, n, dtype=torch.float32, device=DEV)
    loss, grad = net.loss_grad(ls, dev(r, torch.float32), dev(e, torch.float32), zs=zs_d, out=out_d)
    zs = zs_d.cpu().numpy().astype(np.float64)
    # the SGS feature against the oracle's multicolor SGS (same coloring)
    cp, cr = ls.coloring()
    for b in range(B):
        want = oracle.mcsgs_apply(rp, ci, v, cp, cr, r[b].astype(np.float64))
        assert rel(zs[b], want) <= 1e-5
    out_ref, cache = G.forward(th.astype(np.float64), gr, r.astype(np.float64), zs, rnd=True)
    L_ref, _ = G.loss(out_ref, gr, r.astype(np.float64), e.astype(np.float64))
    g_ref = G.backward(th.astype(np.float64), gr, cache, out_ref, r.astype(np.float64), e.astype(np.float64), rnd=True)
    out = out_d.cpu().numpy()
    print("  out rel %.2e  loss %.6f vs %.6f" % (rel(out, out_ref), loss, L_ref), flush=True)
    assert rel(out, out_ref) <= 5e-3
    assert abs(loss - L_ref) <= 2e-3 * max(1.0, abs(L_ref))
    blocks = [("w_in", o["w_in"], o["b_in"]), ("b_in", o["b_in"], o["w0"])]
    blocks += [(f"w{l}", o[f"w{l}"], o[f"b{l}"]) for l in range(4)]
    blocks += [(f"b{l}", o[f"b{l}"], o[f"b{l}"] + 32) for l in range(4)]
    blocks += [("w_out", o["w_out"], o["c"]), ("c", o["c"], o["total"])]
    for name, a, b in blocks:
        err = rel(grad[a:b].astype(np.float64), g_ref[a:b])
        print("  grad %-6s rel %.2e |g| %.2e" % (name, err, np.linalg.norm(g_ref[a:b])), flush=True)
        assert err <= 2e-2, name


@pytest.mark.parametrize("n,ids,runtime_width", [(7, False, False), (7, False, True), (17, True, False)])
def test_inference_path_matches_oracle(oracle, monkeypatch, n, ids, runtime_width):
    """The solver's inference kernels (message tiles in LDS, BF16 dot2 aggregation,
    transposed MFMA layers, fused output) against the oracle's forward pass with
    the same weights; explicit shuffled node ids and the runtime-width kernel too."""
    print("case: learned inference vs oracle, n=%d ids=%s runtime_width=%s" % (n, ids, runtime_width), flush=True)
    monkeypatch.setenv("NALU_GNN_RUNTIME_WIDTH", "1" if runtime_width else "0")
    if ids:
        P = lowmach_problem(n, rho_sigma=0.3, seed=4321)
        nid = np.random.default_rng(6).permutation(P["coords"].shape[0]).astype(np.int64) + 1
        mesh = na.Mesh(P["coords"], P["conn"], nid)
        ls = na.LinearSystem(1, na.SolverConfig("fgmres", "learned", precision="mixed"))
        ls.buildElemToNodeGraph(mesh)
        ls.finalizeLinearSystem()
        ls.zeroSystem()
        na.assemble_continuity(ls, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), tau=1.0,
                               mode="gather")
        ls.loadComplete()
        ls.precond_setup()
    else:
        _, _, ls = poisson_system(n)
    rp, ci, v, _ = ls.csr()
    n = len(rp) - 1
    gr = G.Graph(rp, ci, v)
    net = na.LearnedPreconditioner(num_layers=4, seed=9)
    o = G.layout(4)
    rng = np.random.default_rng(8)
    th = net.get_params().astype(np.float64)
    th[o["w_out"]:o["c"]] = rng.normal(0, 0.3, 32)
    th[o["c"]] = 0.2
    for l in range(4):
        th[o[f"b{l}"]:o[f"b{l}"] + 32] = rng.normal(0, 0.1, 32)
    th = th.astype(np.float32)
    net.set_params(th)
    r = rng.standard_normal(n).astype(np.float32)
    z = torch.zeros(n, dtype=torch.float32, device=DEV)
    net.apply(ls, dev(r, torch.float32), z)
    zs = torch.zeros(1, n, dtype=torch.float32, device=DEV)
    net.loss_grad(ls, dev(r[None], torch.float32), dev(r[None], torch.float32), zs=zs)
    want = G.apply(th.astype(np.float64), gr, r[None].astype(np.float64), zs.cpu().numpy().astype(np.float64))[0]
    err = rel(z.cpu().numpy().astype(np.float64), want)
    print("  z rel %.2e" % err, flush=True)
    assert err <= 5e-3


def test_untrained_network_is_sgs():
    print("case: untrained network == SGS, n=8", flush=True)
    _, _, ls = poisson_system(8)
    n = ls.sizes()[0]
    net = na.LearnedPreconditioner(seed=1)
    r = dev(np.random.default_rng(2).standard_normal(n), torch.float32)
    z = torch.zeros_like(r)
    net.apply(ls, r, z)
    zs = torch.zeros(1, n, dtype=torch.float32, device=DEV)
    net.loss_grad(ls, r[None].contiguous(), r[None].contiguous(), zs=zs)
    # a pure-Neumann system: the output is projected off the constants
    zs0 = zs[0].cpu().numpy().astype(np.float64)
    assert rel(z.cpu().numpy(), zs0 - zs0.mean()) <= 1e-6


def amg_base_system(n, method="fgmres", prec="learned"):
    P = lowmach_problem(n, rho_sigma=0.3, seed=4321)
    mesh = na.Mesh(P["coords"], P["conn"])
    ls = na.LinearSystem(1, na.SolverConfig(method, prec, tolerance=1e-8, max_iterations=2000, kspace=30,
                                            precision="mixed"))
    ls.set_muelu_params(na.MueLuParams(max_coarse_size=30, cheb_degree=2))
    if prec == "learned":
        ls.set_learned_base("muelu")
    ls.buildElemToNodeGraph(mesh)
    ls.finalizeLinearSystem()
    ls.zeroSystem()
    na.assemble_continuity(ls, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), tau=1.0,
                           mode="gather")
    ls.loadComplete()
    ls.removeRhsMean()
    ls.precond_setup()
    return P, mesh, ls


def test_untrained_network_is_the_amg_base():
    """With the AMG base the network's feature is the FP32 V-cycle of r (within
    FP32 rounding of the FP64 V-cycle of a muelu system), the untrained network
    returns it, and a non-finite network falls back to the V-cycle."""
    print("case: learned preconditioner on the AMG base, n=10", flush=True)
    P, mesh, ls = amg_base_system(10)
    _, _, lm = amg_base_system(10, prec="muelu")
    n = ls.sizes()[0]
    net = na.LearnedPreconditioner(seed=1)
    r64 = np.random.default_rng(3).standard_normal(n)
    r64 -= r64.mean()
    r = dev(r64, torch.float32)
    z = torch.zeros_like(r)
    net.apply(ls, r, z)
    zs = torch.zeros(1, n, dtype=torch.float32, device=DEV)
    net.loss_grad(ls, r[None].contiguous(), r[None].contiguous(), zs=zs)
    zamg = torch.zeros(n, dtype=torch.float64, device=DEV)
    lm.precond_apply(dev(r.double().cpu().numpy()), zamg)
    z_np, zs_np, za = z.cpu().numpy(), zs[0].cpu().numpy().astype(np.float64), zamg.cpu().numpy()
    zp = zs_np - zs_np.mean()  # pure-Neumann system: the output is projected off the constants
    print("  |net - feature| %.2e, |feature - FP64 V-cycle| %.2e" % (rel(z_np, zp), rel(zs_np, za)), flush=True)
    assert rel(z_np, zp) <= 1e-6
    assert rel(zs_np, za) <= 1e-4
    th = net.get_params()
    th[:] = np.nan
    net.set_params(th)
    ls.set_learned_preconditioner(net)
    b = dev(poisson_rhs(n, 5))
    x = torch.zeros(n, dtype=torch.float64, device=DEV)
    st, it, rr = ls.solve_rhs(b, x)
    print("  NaN network: status %d iters %d resid %.2e fallback %s" % (st, it, rr, ls.learned_fallback()),
          flush=True)
    assert st == 0 and rr <= 1e-8 and ls.learned_fallback() and it <= 60


def test_training_on_the_amg_base():
    """Pairs harvested with the AMG base store the V-cycle feature; a network
    trained on them solves inside FGMRES without falling back."""
    print("case: train on the AMG base, n=12", flush=True)
    P, mesh, ls = amg_base_system(12)
    n = ls.sizes()[0]
    data = na.HarvestedPairs(ls, num_rhs=12, pairs_per_rhs=10, seed=11, tol=1e-10)
    net = na.LearnedPreconditioner(seed=2)
    hist = net.train(ls, data, epochs=6, batch=8, seed=1)
    print("  epoch losses", np.round(hist, 4).tolist(), flush=True)
    assert np.all(np.isfinite(hist)) and hist[-1] <= hist[0]
    b = dev(poisson_rhs(n, 99))
    ls.set_learned_preconditioner(net)
    x = torch.zeros(n, dtype=torch.float64, device=DEV)
    st1, it1, r1 = ls.solve_rhs(b, x)
    _, _, lm = amg_base_system(12, prec="muelu")
    x.zero_()
    st0, it0, r0 = lm.solve_rhs(b, x)
    print("  FGMRES iterations: AMG %d, learned on AMG %d (fallback %s)" % (it0, it1, ls.learned_fallback()),
          flush=True)
    assert st0 == 0 and st1 == 0 and r1 <= 1e-8 and not ls.learned_fallback()


def test_harvest_pairs_are_consistent(oracle):
    print("case: harvest n=8", flush=True)
    _, _, ls = poisson_system(8, na.SolverConfig("cg", "sgs"))  # plain FP64 system, as the pairs are defined
    rp, ci, v, _ = ls.csr()
    n = len(rp) - 1
    data = na.HarvestedPairs(ls, num_rhs=3, pairs_per_rhs=5, seed=7, tol=1e-10)
    assert len(data) == 15
    its = data.cg_iterations()
    assert all(5 <= k < 500 for k in its)
    r = torch.zeros(n, dtype=torch.float32, device=DEV)
    e = torch.zeros_like(r)
    for idx in (0, 4, 9, 14):
        data.pair(idx, r, e)
        rh, eh = r.cpu().numpy().astype(np.float64), e.cpu().numpy().astype(np.float64)
        assert abs(np.linalg.norm(rh) - 1.0) < 1e-6
        # pairs are kept only where the converged solution fixes e to 1e-3 of |e|
        assert rel(oracle.spmv(rp, ci, v, eh), rh) <= 2e-3, idx
        assert float(eh @ rh) > 0


def test_set_pair_stores_a_caller_pair(oracle):
    """A caller's (r = A e, e) is stored like a harvested pair: scaled by 1/|r|,
    e off its mean; the stored feature is the base preconditioner's output."""
    print("case: set_pair n=8", flush=True)
    _, _, ls = poisson_system(8, na.SolverConfig("cg", "sgs"))
    rp, ci, v, _ = ls.csr()
    n = len(rp) - 1
    data = na.HarvestedPairs(ls, num_rhs=2, pairs_per_rhs=2, seed=7, tol=1e-10)
    e = np.random.default_rng(3).standard_normal(n) + 5.0  # mean far from zero
    r = oracle.spmv(rp, ci, v, e)
    data.set_pair(2, dev(r), dev(e))
    rg = torch.zeros(n, dtype=torch.float32, device=DEV)
    eg = torch.zeros_like(rg)
    data.pair(2, rg, eg)
    nr = np.linalg.norm(r)
    assert rel(rg.cpu().numpy().astype(np.float64), r / nr) <= 1e-6
    assert rel(eg.cpu().numpy().astype(np.float64), (e - e.mean()) / nr) <= 1e-6
    # the other pairs are untouched and the set trains
    data.pair(0, rg, eg)
    assert abs(np.linalg.norm(rg.cpu().numpy().astype(np.float64)) - 1.0) < 1e-6
    net = na.LearnedPreconditioner(seed=2)
    hist = net.train(ls, data, epochs=2, batch=2, lr=1e-3, seed=1)
    assert np.all(np.isfinite(hist))
    with pytest.raises(na.NaluError, match="index out of range"):
        data.set_pair(4, dev(r), dev(e))
    with pytest.raises(na.NaluError, match="zero or non-finite"):
        data.set_pair(1, dev(np.zeros(n)), dev(e))


def test_self_pair_rounds():
    """The training recipe's self-pair rounds replace pairs with the network's
    own FGMRES residuals / errors and retrain with validation."""
    print("case: self-pair rounds n=10", flush=True)
    from nalu_amd.training import train_pressure_preconditioner
    P = lowmach_problem(10, rho_sigma=0.3, seed=4321)
    net, rec = train_pressure_preconditioner(P, num_rhs=16, pairs_per_rhs=4, harvest_tol=1e-6, epochs=2, batch=8,
                                             lr=3e-3, device=DEV, self_pair_rounds=1, self_pair_fraction=0.5,
                                             self_pair_caps=(2, 5), self_pair_epochs=2)
    rounds = rec["self_pair_rounds"]
    assert len(rounds) == 1 and rounds[0]["pairs_replaced"] == 32
    # the first training's record is kept apart from the round's
    assert len(rec["validation_fgmres_iters_per_epoch"]) == 2 and 0 <= rec["kept_epoch"] < 2
    assert rec["validation_fgmres_iters_per_epoch"] is not rounds[0]["validation_fgmres_iters_per_epoch"]
    assert len(rounds[0]["validation_fgmres_iters_per_epoch"]) == 2
    assert all(np.isfinite(rounds[0]["epoch_loss"]))
    assert 0 <= rounds[0]["kept_epoch"] < 2


def test_training_gains_over_sgs():
    print("case: train n=12 and FGMRES learned vs SGS", flush=True)
    cfg = na.SolverConfig("fgmres", "learned", tolerance=1e-8, max_iterations=2000, kspace=30, precision="mixed")
    P, mesh, ls = poisson_system(12, cfg)
    n = ls.sizes()[0]
    data = na.HarvestedPairs(ls, num_rhs=12, pairs_per_rhs=20, seed=11, tol=1e-10)
    net = na.LearnedPreconditioner(seed=2)
    hist = net.train(ls, data, epochs=12, batch=8, lr=3e-3, seed=1)
    print("  epoch losses", np.round(hist, 4).tolist(), flush=True)
    assert np.all(np.isfinite(hist)) and hist[-1] < 0.7 * hist[0]
    b = dev(poisson_rhs(n, 99))

    def solve(prec):
        ls_s = na.LinearSystem(1, na.SolverConfig("fgmres", prec, tolerance=1e-8, max_iterations=2000, kspace=30,
                                                  precision="mixed"))
        ls_s.buildElemToNodeGraph(mesh)
        ls_s.finalizeLinearSystem()
        ls_s.zeroSystem()
        na.assemble_continuity(ls_s, mesh, dev(P["rho"]), dev(P["p"]), dev(P["u"]), dev(P["gpdx"]), tau=1.0,
                               mode="gather")
        ls_s.loadComplete()
        if prec == "learned":
            ls_s.set_learned_preconditioner(net)
        x = torch.zeros(n, dtype=torch.float64, device=DEV)
        st, it, rr = ls_s.solve_rhs(b, x)
        return st, it, rr, (ls_s.learned_fallback() if prec == "learned" else False)

    st0, it0, r0, _ = solve("sgs")
    st1, it1, r1, fb = solve("learned")
    print("  FGMRES iterations: sgs %d, learned %d (fallback %s); resid %.2e %.2e" % (it0, it1, fb, r0, r1),
          flush=True)
    assert st0 == 0 and st1 == 0 and r1 <= 1e-8
    assert not fb
    assert it1 <= 0.8 * it0


def test_nonfinite_network_falls_back_to_sgs():
    print("case: NaN network falls back, n=8", flush=True)
    P, mesh, ls = poisson_system(8)
    n = ls.sizes()[0]
    net = na.LearnedPreconditioner(seed=1)
    th = net.get_params()
    th[:] = np.nan
    net.set_params(th)
    ls.set_learned_preconditioner(net)
    b = dev(poisson_rhs(n, 5))
    x = torch.zeros(n, dtype=torch.float64, device=DEV)
    st, it, rr = ls.solve_rhs(b, x)
    print("  status %d iters %d resid %.2e fallback %s" % (st, it, rr, ls.learned_fallback()), flush=True)
    assert st == 0 and rr <= 1e-8 and ls.learned_fallback()


def test_epoch_ranges_and_validation_selection():
    """Training split into single epochs (with a validation call in between)
    follows the one-call run exactly; validation keeps the best epoch."""
    print("case: epoch-range training and validation selection, n=8", flush=True)
    cfg = na.SolverConfig("fgmres", "learned", tolerance=1e-8, max_iterations=400, kspace=30, precision="mixed")
    P, mesh, ls = poisson_system(8, cfg)
    data = na.HarvestedPairs(ls, num_rhs=8, pairs_per_rhs=4, seed=3, tol=1e-8)
    a = na.LearnedPreconditioner(seed=5)
    ha = a.train(ls, data, epochs=3, batch=8, lr=3e-3, seed=2)
    b = na.LearnedPreconditioner(seed=5)
    calls = []

    def last_is_best(net):
        calls.append(1)
        return -len(calls)

    hb = b.train(ls, data, epochs=3, batch=8, lr=3e-3, seed=2, validate=last_is_best)
    np.testing.assert_array_equal(ha, hb)
    np.testing.assert_array_equal(a.get_params(), b.get_params())
    assert b.best_epoch == 2 and len(b.validation) == 3
    # real validation: FGMRES iterations on the system itself; the kept network scores the minimum
    n = ls.sizes()[0]
    bb = dev(poisson_rhs(n, 17))
    x = torch.zeros(n, dtype=torch.float64, device=DEV)

    def iters(net):
        ls.set_learned_preconditioner(net)
        x.zero_()
        st, it, rr = ls.solve_rhs(bb, x)
        return it + (10000 if ls.learned_fallback() else 0)

    c = na.LearnedPreconditioner(seed=6)
    c.train(ls, data, epochs=4, batch=8, lr=3e-3, seed=4, validate=iters)
    print("  validation", c.validation, "kept", c.best_epoch, flush=True)
    assert iters(c) == min(c.validation)
