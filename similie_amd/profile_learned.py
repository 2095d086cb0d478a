"""Fixed workloads of the learned preconditioner for timing and profiling.

    python -m similie_amd.profile_learned infer N     N inference passes of the trained
                                                      network on 64 random 256^2 inputs
    python -m similie_amd.profile_learned forward N   the same with the stored-activation pass
    python -m similie_amd.profile_learned solve METHOD one warm-up and two timed solves of the
                                                      64 held-out problems (fcg_learned,
                                                      fcg_learned_mixed, cg, cg_mixed, cg_plain)

Each prints one line per timed unit; under rocprofv3 the kernels of the
last passes are the ones to read.
"""
import sys

import numpy as np
import torch

from similie_amd import learned as Lr
from similie_amd.preconditioner_training import DEFAULT_WEIGHTS, make_problems


def trained_network(batch=64, n=256):
    net = Lr.UNet(batch, n, n)
    net.set_params(torch.from_numpy(np.load(DEFAULT_WEIGHTS)))
    return net


def passes(mode, count):
    net = trained_network()
    x = torch.randn((64, 256, 256, 2), dtype=torch.float32, device="cuda")
    run = net.forward if mode == "forward" else net.infer
    for _ in range(3):
        run(x)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(count):
        run(x)
    stop.record()
    torch.cuda.synchronize()
    print(f"{mode}: {start.elapsed_time(stop) / count * 1e3:.1f} us per pass", flush=True)


def solves(method):
    kappa, b = make_problems(64, 256, 1.5, 16.0, 900000)
    pb = Lr.PoissonBatch(kappa)
    net = trained_network() if method.startswith("fcg") else None
    pb.solve(b, method=method, tol=1e-10, net=net, max_iterations=20)
    for k in range(2):
        _, res, secs = pb.solve(b, method=method, tol=1e-10, net=net)
        its = [r["iterations"] for r in res]
        print(f"{method} {k}: {secs * 1e3:.2f} ms, iterations mean {np.mean(its):.2f} max {max(its)}", flush=True)


def main(argv):
    if len(argv) != 2 or argv[0] not in ("infer", "forward", "solve"):
        print(__doc__, file=sys.stderr)
        return 2
    if argv[0] == "solve":
        solves(argv[1])
    else:
        passes(argv[0], int(argv[1]))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
