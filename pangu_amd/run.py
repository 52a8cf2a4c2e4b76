"""Run one of the standard configurations on the GPU and report its rate.

    python -m pangu_amd.run fm_torus_cks --cycles 20
    python -m pangu_amd.run torus3d --cycles 100 --mode graph --output out.phdf

The counterpart of running the reference executable on an input file
(pangu/problem/*/inputfile): the configuration functions of
pangu_amd.configs stand in for the input files.  Prints one JSON line with
the block count, zones, ms per cycle (HIP events around the timed cycles)
and zone-cycles/s; --output also writes the final state as .phdf + .xdmf.
"""
import argparse
import json
import time

from . import configs

CONFIGS = {
    "orszag_tang": configs.orszag_tang,
    "torus2d": configs.torus2d,
    "torus3d": configs.torus3d,
    "fm_torus_cks": configs.fm_torus_cks,
}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("config", choices=sorted(CONFIGS))
    ap.add_argument("--cycles", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", default="graph", choices=["graph", "fused", "tasks"])
    ap.add_argument("--output", default=None, help="write the final state to this .phdf path")
    args = ap.parse_args(argv)
    from .sim import HipSim
    cfg = CONFIGS[args.config]()
    t0 = time.perf_counter()
    sim = HipSim(cfg)
    sim.init()
    t_init = time.perf_counter() - t0
    per_block = [(sim.ends[b][d] - sim.starts[b][d] + 1) for b in range(sim.nblocks) for d in range(3)]
    zones = sum(per_block[3 * b] * per_block[3 * b + 1] * per_block[3 * b + 2] for b in range(sim.nblocks))
    sim.step(args.warmup, args.mode)
    ms = sim.timed_cycles(args.cycles, args.mode)
    out = {"config": args.config, "blocks": sim.nblocks, "zones": zones, "cycles": args.cycles,
           "mode": args.mode, "ms_per_cycle": ms / args.cycles,
           "zone_cycles_per_s": zones * args.cycles / (ms * 1e-3), "init_s": t_init,
           "time": sim.time, "dt": sim.dt}
    if args.output:
        from . import phdf
        out["output"] = phdf.write(sim, args.output, ncycle=args.warmup + args.cycles)
    sim.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
