"""Problem generator, training-data generation, training and evaluation of
the learned preconditioner (no reference counterpart).

Problems: Dirichlet-zero Poisson problems -div(κ grad u) = b on n x n
unit-spaced grids, κ = exp(g) with g a Gaussian random field (squared-
exponential covariance, standard deviation σ, correlation length ℓ cells),
b ~ N(0, 1) on interior nodes.

Training data are (residual, error) pairs that plain CG itself produces:
Jacobi CG runs every training problem to 1e-10 while the batch solver
records its iterates x_k; the error of record k is e_k = x* − x_k with x*
the converged solution.  The solver applies the network inside a two-level
cycle, after a damped-Jacobi pre-smoothing step, so each pair is taken after
that step (see generate_pairs).  The network maps [r/rms(r), normalised
log κ] to a correction direction and is trained with the direction loss
1 − cos²_A(e, N) (sil_unet_train_step).

    python -m similie_amd.preconditioner_training --out similie_amd/weights/unet_poisson256.npy

The random fields and right-hand sides are synthetic inputs built with torch
(device memory plumbing); every solve and every network pass runs through
libsil_hip.so.
"""
import argparse
import json
import math
import time
from pathlib import Path

import numpy as np
import torch

from . import learned as Lr

DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "weights" / "unet_poisson256.npy"


def gaussian_random_field(batch, n, sigma, corr, seed, device="cuda"):
    """[batch][n][n] fields with covariance σ² exp(−r²/(2 ℓ²)) (spectral
    synthesis on a periodic 2n grid, cropped), each rescaled to mean 0 and
    standard deviation σ."""
    gen = torch.Generator(device=device).manual_seed(seed)
    m = 2 * n
    white = torch.randn((batch, m, m), generator=gen, dtype=torch.float64, device=device)
    k = torch.fft.fftfreq(m, d=1.0, device=device, dtype=torch.float64) * 2 * math.pi
    k2 = k[:, None] ** 2 + k[None, :] ** 2
    amp = torch.exp(-0.25 * k2 * corr * corr)  # sqrt of the squared-exponential spectrum
    g = torch.fft.ifft2(torch.fft.fft2(white) * amp).real[:, :n, :n]
    g = g - g.mean(dim=(1, 2), keepdim=True)
    g = g / g.std(dim=(1, 2), keepdim=True)
    return sigma * g


def make_problems(batch, n, sigma, corr, seed, device="cuda"):
    """(κ, b) of `batch` problems; seeds of different problem sets must differ."""
    kappa = torch.exp(gaussian_random_field(batch, n, sigma, corr, seed, device))
    gen = torch.Generator(device=device).manual_seed(seed + 7919)
    b = torch.randn((batch, n, n), generator=gen, dtype=torch.float64, device=device)
    b[:, 0, :] = 0
    b[:, -1, :] = 0
    b[:, :, 0] = 0
    b[:, :, -1] = 0
    return kappa, b


OMEGA = 2.0 / 3.0  # damping of the Jacobi smoothing in the solver's two-level preconditioner


def jacobi_inverse(kappa):
    """D⁻¹ of the batch operator (interior nodes; 0 on the boundary ring), FP64."""
    wx = torch.empty_like(kappa)
    wy = torch.empty_like(kappa)
    wx[:, :-1, :] = 0.5 * (kappa[:, :-1, :] + kappa[:, 1:, :])
    wx[:, -1, :] = kappa[:, -1, :]
    wy[:, :, :-1] = 0.5 * (kappa[:, :, :-1] + kappa[:, :, 1:])
    wy[:, :, -1] = kappa[:, :, -1]
    d = torch.zeros_like(kappa)
    d[:, 1:-1, 1:-1] = 1.0 / (wx[:, 1:-1, 1:-1] + wx[:, :-2, 1:-1] + wy[:, 1:-1, 1:-1] + wy[:, 1:-1, :-2])
    return d


def generate_pairs(n_problems, n, sigma, corr, seed, pairs_per_problem, batch=64, record_every=8, max_records=400,
                   net=None, max_iterations=20000, log=print):
    """Training pairs of the network as the solver uses it.

    Jacobi CG solves every problem to 1e-10 while recording its iterates
    x_k (with `net`, flexible CG with that network records instead, and a
    separate CG solve provides x*).  For each picked record the pair is
    formed after the solver's pre-smoothing step: r = b − A x_k,
    e = x* − x_k, s = ω D⁻¹ r, r₁ = r − A s, e₁ = e − s (so A e₁ = r₁), both
    scaled by 1/rms(r₁).  Returns (inputs [P][n][n][2], errors [P][n][n],
    wx, wy, problem index)."""
    inputs, errors, wxs, wys, owners = [], [], [], [], []
    for start in range(0, n_problems, batch):
        B = min(batch, n_problems - start)
        kappa, b = make_problems(B, n, sigma, corr, seed + start)
        if not bool(torch.isfinite(kappa).all()):
            raise RuntimeError(f"non-finite coefficients for seed {seed + start}")
        pb = Lr.PoissonBatch(kappa)
        t0 = time.time()
        if net is None:
            x, res, secs, (rr, xr) = pb.solve(b, method="cg", tol=1e-10, record_every=record_every,
                                              max_records=max_records)
            ok = np.array([r["converged"] for r in res])
        else:
            x, res0, _ = pb.solve(b, method="cg", tol=1e-10)
            ok = np.array([r["converged"] for r in res0])
            _, res, secs, (rr, xr) = pb.solve(b, method="fcg_learned", tol=1e-10, net=net,
                                              record_every=record_every, max_records=max_records,
                                              max_iterations=max_iterations)
        its = np.array([r["iterations"] for r in res])
        log(f"pairs: problems {start}..{start + B - 1}: CG iterations {its.min()}..{its.max()}, "
            f"{int(ok.sum())}/{B} converged, {xr.shape[0]} records, {time.time() - t0:.1f} s")
        wx, wy, lk = pb.weights()
        dinv = jacobi_inverse(kappa)
        nrec = xr.shape[0]
        pick = np.zeros((nrec, B), dtype=bool)
        for p in range(B):
            if not ok[p]:
                continue
            valid = int(min(nrec, max(1, its[p] // record_every)))
            pick[np.unique(np.linspace(0, valid - 1, pairs_per_problem).round().astype(int)), p] = True
        for k in range(nrec):
            sel = torch.from_numpy(np.nonzero(pick[k])[0]).cuda()
            if sel.numel() == 0:
                continue
            xk = xr[k]
            r = b - pb.apply(xk)
            s = OMEGA * dinv * r
            r1 = r - pb.apply(s.contiguous())
            e1 = (x - xk) - s
            rho = torch.sqrt(torch.mean(r1 * r1, dim=(1, 2), keepdim=True)).clamp_min(1e-300)
            inputs.append(torch.stack([(r1 / rho).float()[sel], lk[sel]], dim=-1))
            errors.append((e1 / rho).float()[sel])
            owners.append(sel + start)
        wxs.append(wx)
        wys.append(wy)
        del rr, xr
    return (torch.cat(inputs), torch.cat(errors), torch.cat(wxs), torch.cat(wys), torch.cat(owners))


def train(net, data, epochs, batch=32, lr_max=1e-3, lr_min=1e-5, seed=0, log=print):
    inputs, errors, wx, wy, owners = data
    P = inputs.shape[0]
    steps = epochs * (P // batch)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    step = 0
    history = []
    t0 = time.time()
    for ep in range(epochs):
        perm = torch.randperm(P, generator=gen, device="cuda")
        losses = []
        for s in range(P // batch):
            idx = perm[s * batch: (s + 1) * batch]
            lr = lr_min + 0.5 * (lr_max - lr_min) * (1 + math.cos(math.pi * step / max(1, steps)))
            own = owners[idx]
            loss = net.train_step(inputs[idx].contiguous(), errors[idx].contiguous(), wx[own].contiguous(),
                                  wy[own].contiguous(), lr)
            losses.append(loss.mean())
            step += 1
        mean = torch.stack(losses).mean().item()
        history.append(mean)
        log(f"epoch {ep}: mean direction loss {mean:.4f} ({time.time() - t0:.0f} s)")
    return history


def evaluate(net, n_problems, n, sigma, corr, seed,
             methods=("cg_plain", "cg", "cg_mixed", "fcg_learned", "fcg_learned_mixed"),
             tol=1e-10, ablation=True, log=print):
    """Time to an FP64-verified tol of each method on one batch of held-out
    problems.  With ablation, flexible CG also runs with a network whose
    parameters are all zero: the same two-level cycle without the learned
    correction."""
    kappa, b = make_problems(n_problems, n, sigma, corr, seed)
    pb = Lr.PoissonBatch(kappa)
    runs = [(m, net if m.startswith("fcg") else None) for m in methods]
    if ablation:
        zero = Lr.UNet(net.max_batch, n, n)
        zero.set_params(torch.zeros(Lr.param_count()))
        runs.append(("fcg_without_network", zero))
    out = {}
    for name, nn in runs:
        m = "fcg_learned" if name == "fcg_without_network" else name
        # first call warms up (graph instantiation, allocation), second is timed
        pb.solve(b, method=m, tol=tol, net=nn, max_iterations=20)
        x, res, secs = pb.solve(b, method=m, tol=tol, net=nn)
        its = [r["iterations"] for r in res]
        worst = max(r["true_relative_residual"] for r in res)
        out[name] = {"seconds": secs, "iterations_mean": float(np.mean(its)), "iterations_max": int(max(its)),
                     "converged": int(sum(r["converged"] for r in res)), "worst_true_relative_residual": worst}
        log(f"{name}: {secs * 1e3:.1f} ms, iterations mean {np.mean(its):.1f} max {max(its)}, "
            f"worst residual {worst:.2e}")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--sigma", type=float, default=1.5)
    ap.add_argument("--corr", type=float, default=16.0)
    ap.add_argument("--problems", type=int, default=200)
    ap.add_argument("--pairs-per-problem", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--eval-seed", type=int, default=900000)
    ap.add_argument("--eval-problems", type=int, default=64)
    ap.add_argument("--solver-rounds", type=int, default=0,
                    help="extra rounds on pairs recorded along flexible CG with the current network")
    ap.add_argument("--solver-epochs", type=int, default=4)
    ap.add_argument("--init", type=Path, default=None, help="start from these parameters (.npy)")
    ap.add_argument("--out", type=Path, default=DEFAULT_WEIGHTS)
    ap.add_argument("--log", type=Path, default=None)
    a = ap.parse_args(argv)
    logf = open(a.log, "w") if a.log else None

    def log(msg):
        print(msg, flush=True)
        if logf:
            logf.write(msg + "\n")
            logf.flush()

    t0 = time.time()
    data = generate_pairs(a.problems, a.n, a.sigma, a.corr, a.seed, a.pairs_per_problem, log=log)
    log(f"{data[0].shape[0]} pairs in {time.time() - t0:.0f} s")
    net = Lr.UNet(max(a.batch, a.eval_problems, 64), a.n, a.n, seed=a.seed)
    if a.init is not None:
        net.set_params(torch.from_numpy(np.load(a.init)))
        log(f"initial parameters from {a.init}")
    hist = train(net, data, a.epochs, batch=a.batch, lr_max=a.lr, log=log) if a.epochs > 0 else []
    # optional further rounds on the residuals flexible CG itself meets
    for rnd in range(a.solver_rounds):
        t0 = time.time()
        more = generate_pairs(a.problems, a.n, a.sigma, a.corr, a.seed, a.pairs_per_problem // 2, net=net,
                              record_every=4, max_records=300, max_iterations=1200, log=log)
        data = tuple(torch.cat([u, v]) if k != 2 and k != 3 else u for k, (u, v) in enumerate(zip(data, more)))
        log(f"round {rnd + 1}: {more[0].shape[0]} pairs from flexible CG ({time.time() - t0:.0f} s), "
            f"{data[0].shape[0]} in total")
        del more
        hist += train(net, data, a.solver_epochs, batch=a.batch, lr_max=a.lr / 3, log=log)
    del data
    params = net.get_params().cpu().numpy()
    a.out.parent.mkdir(parents=True, exist_ok=True)
    np.save(a.out, params)
    log(f"saved {a.out} ({params.size} parameters)")
    res = evaluate(net, a.eval_problems, a.n, a.sigma, a.corr, a.eval_seed, log=log)
    log(json.dumps({"loss_history": hist, "evaluation": res}))


if __name__ == "__main__":
    main()
