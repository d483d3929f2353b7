"""Per-level timing of one registration, as the serving and benchmark workloads
run it: for each level of the shrink / smoothing schedule, the metric's
set-up (smoothing, gradient filter, tables), the optimizer's iterations and
one evaluation's device time, with the RegularStep loop on the device (the
default) or on the host.

    python -m itk_amd.level_timing [--kind affine|bspline] [--size 256] [--mesh 8] [--host-loop]
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from . import synth
from .batch import shared_evaluator
from .metric import MattesMutualInformationImageToImageMetricv4 as Metric
from .serving import RegistrationRequest, build_registration


def time_registration(kind: str, size: int, mesh: int = 8, device_loop: bool = True, seed: int = 11, log=print):
    shared_evaluator().set_step_loop(device_loop, 8)
    shape = (size, size, size)
    p = synth.make_affine_problem(shape, seed) if kind == "affine" else \
        synth.make_bspline_problem(shape, seed, mesh=8, sigma_vox=2.0)
    out = []
    for rep in range(2):  # the first run pays the one-time costs (module load, first allocations)
        req = RegistrationRequest(0, kind, p.fixed, p.moving, p.geometry, mesh=mesh,
                                  iterations=100 if kind == "affine" else 50)
        m = Metric()
        init_times = []
        orig = m.Initialize

        def timed_init(orig=orig, init_times=init_times):
            t = time.perf_counter()
            orig()
            init_times.append(time.perf_counter() - t)

        m.Initialize = timed_init
        reg, tf, opt = build_registration(req, m, native_loop=True)
        t0 = time.perf_counter()
        reg.Update()
        total = time.perf_counter() - t0
        levels = []
        for lv, ti in zip(reg.level_log, init_times):
            m.SetVirtualDomain(p.geometry.shrink(lv["shrink"]))
            sig = reg._level_sigmas(lv["sigma"], p.geometry)
            m.SetSmoothingSigmas(sig, sig)
            orig()
            dev = []
            for _ in range(6):
                m.GetValueAndDerivative()
                dev.append(m.GetLastTiming()["total_ms"])
            its = max(1, lv["iterations"])
            levels.append({"level": lv["level"], "virtual_size": list(lv["virtual_size"]), "samples": lv["samples"],
                           "init_ms": ti * 1e3, "iterations": lv["iterations"],
                           "iterate_ms": (lv["seconds"] - ti) * 1e3,
                           "ms_per_iteration": (lv["seconds"] - ti) * 1e3 / its,
                           "device_ms_per_eval": float(np.median(dev[1:]))})
        m.close()
        out = {"kind": kind, "size": size, "mesh": mesh, "device_loop": device_loop, "total_ms": total * 1e3,
               "levels": levels}
    log("== %s %d^3%s, RegularStep loop on the %s: total %.1f ms" % (
        kind, size, " mesh %d" % mesh if kind == "bspline" else "", "device" if device_loop else "host",
        out["total_ms"]))
    for lv in out["levels"]:
        log("  level %d virtual %s (%d samples): set-up %.2f ms, %d iterations in %.2f ms (%.3f ms each), "
            "device %.3f ms per evaluation" % (lv["level"], "x".join(map(str, lv["virtual_size"])), lv["samples"],
                                               lv["init_ms"], lv["iterations"], lv["iterate_ms"],
                                               lv["ms_per_iteration"], lv["device_ms_per_eval"]))
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--kind", choices=("affine", "bspline"), default="affine")
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--mesh", type=int, default=8)
    ap.add_argument("--host-loop", action="store_true", help="RegularStep on the host (comparison)")
    args = ap.parse_args(argv)
    time_registration(args.kind, args.size, args.mesh, not args.host_loop, log=lambda s: print(s, flush=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
