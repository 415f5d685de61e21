"""Per-row phase breakdown of the alignment kernels, from the row-phase
timing build of the library (make -C spoa_amd phases, which compiles the
kernels with SPOA_ROW_PHASES into spoa_amd/_phases/).

    python -m spoa_amd.row_phases --config 2 [--windows 2048] [--merge host|device]

runs one batch of a benchmark workload untimed (the first run warms up),
then a second run with the counters reset, and prints a markdown table of
where wave 0 of every job spends its shader-clock cycles, per row tile and
per job.  The timestamps (s_memtime) wait for the wave's outstanding scalar
loads, so the timing build runs a few percent slower than the product
build; the table gives shares, not the product's absolute times."""
import argparse
import ctypes
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PHASE_LIB = os.path.join(HERE, "_phases", "libspoa_amd.so")

NAMES = ["query profile", "vertical pass", "  of which waiting for predecessor rows", "local pass",
         "scan", "  of which the wave-total barrier", "fix-up and stores", "  of which the wait before the stores",
         "end cell and boundary hand-off", "  of which the hand-off barrier", "end-cell reduction (per job)",
         "traceback (per job)"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4])
    ap.add_argument("--windows", type=int, default=None)
    ap.add_argument("--seed", type=int, default=20240601)
    ap.add_argument("--merge", choices=["host", "device"], default=None)
    args = ap.parse_args()
    if not os.path.exists(PHASE_LIB):
        sys.exit("row_phases: %s is missing; build it with make -C spoa_amd phases" % PHASE_LIB)
    # before the package loads its library (the first lib() call)
    os.environ["SPOA_AMD_LIBRARY"] = PHASE_LIB
    import spoa_amd
    cfg = spoa_amd.CONFIGS[args.config]
    n_windows = args.windows or {1: 1024, 2: 2048, 3: 4096, 4: 256}[args.config]
    merge = args.merge or ("host" if args.config == 2 else "device")
    windows = spoa_amd.config_windows(args.config, args.seed, n_windows)
    ctx = spoa_amd.BatchContext(*cfg["params"], strand=cfg.get("strand", False), msa=True,
                                resident=merge == "device")
    L = spoa_amd.lib()
    out = (ctypes.c_uint64 * 22)()
    ctx.run(windows)  # warm-up
    devs = [ctx.device_of(h) for h in (0, 1)]
    for d in devs:
        if L.spoa_hip_row_phases(d, out, 22, 1) != 22:
            sys.exit("row_phases: " + L.spoa_hip_last_error().decode())
    ctx.run(windows)
    tot = [0] * 22
    for d in devs:
        L.spoa_hip_row_phases(d, out, 22, 1)
        tot = [a + b for a, b in zip(tot, out)]
    tiles, jobs = tot[12], tot[13]
    row_sum = sum(tot[k] for k in (0, 1, 3, 4, 6, 8))
    print("workload: %s, %d windows, %s merge; %d jobs, %d row tiles (wave 0 of every job)"
          % (cfg["name"], n_windows, merge, jobs, tiles))
    print()
    print("| phase | cycles per row tile | share of the row loop |")
    print("|---|---|---|")
    for k in range(10):
        share = "%.1f%%" % (100.0 * tot[k] / row_sum) if row_sum else "-"
        print("| %s | %.0f | %s |" % (NAMES[k], tot[k] / max(tiles, 1), share))
    print()
    print("| per job | cycles | against the job's row loop |")
    print("|---|---|---|")
    for k in (10, 11):
        print("| %s | %.0f | %.1f%% |" % (NAMES[k], tot[k] / max(jobs, 1), 100.0 * tot[k] / max(row_sum, 1)))
    tb = tot[14:22]
    print()
    print("| traceback part | cycles per job | share of the traceback | count per job | cycles each |")
    print("|---|---|---|---|---|")
    for name, cyc, cnt in (("tile staging", tb[0], tb[1]), ("diagonal runs (steps)", tb[2], tb[3]),
                           ("general steps", tb[4], tb[5]), ("left gap walks", tb[6], None),
                           ("up gap walks", tb[7], None)):
        each = "%.0f" % (cyc / cnt) if cnt else "-"
        per = "%.1f" % (cnt / max(jobs, 1)) if cnt is not None else "-"
        print("| %s | %.0f | %.1f%% | %s | %s |" % (name, cyc / max(jobs, 1), 100.0 * cyc / max(tot[11], 1), per, each))


if __name__ == "__main__":
    main()
