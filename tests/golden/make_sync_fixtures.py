#!/usr/bin/env python3
"""Write the multi-rank fixtures of tests/golden/sync/ from the reference's
own program.

The reference's MPI program (src/main.cpp with the workers, FP32 as the
reference builds) is compiled from the reference sources where they lie by
`make -C oracle refmain` (-> oracle/_ref/mcmpi_main, never committed) and
run as `mpirun -n R mcmpi_main config.yaml sync` on this container's CPU,
once per case below.  Kept per case, under tests/golden/sync/<case>/:

  config.yaml      the input (the reference's flat `key: value` format)
  weights.csv      out/weights.csv (src/worker.cpp:287-313), every rank's
                   tallies in rank order
  out_config.yaml  out/config.yaml (src/worker.cpp:218-241), its hostname
                   line replaced by `hostname: HOSTNAME`
  WA.out           rank 0's layer (src/layer.cpp:363-380)
  summary.json     ranks, the stats.csv header, its rows per rank and each
                   rank's nb_cycles summed over its rows (the timings
                   themselves are not data)

Run in a container that has the reference sources and MPICH
(/opt/conda/bin/mpirun); the tests only read the files written here.
"""
import json
import os
import shutil
import socket
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "sync")
EXE = os.path.join(ROOT, "oracle", "_ref", "mcmpi_main")
MPIRUN = os.environ.get("MPIRUN", "/opt/conda/bin/mpirun")

# float32 values, written as the reference's own config.yaml writes them
X_INI = "7.071067690849304199e-01"
MIN_W = "9.999999960041972003e-13"

CASES = {
    # the reference config.yaml's law and source at 2, 3 and 5 ranks
    "r2_c1000": (2, dict(nb_cells=1000, nb_particles=20000, x_ini=X_INI,
                         particle_min_weight=MIN_W, nb_particles_per_cycle=500)),
    "r3_c1000": (3, dict(nb_cells=1000, nb_particles=20000, x_ini=X_INI,
                         particle_min_weight=MIN_W, nb_particles_per_cycle=500)),
    "r5_c1000": (5, dict(nb_cells=1000, nb_particles=20000, x_ini=X_INI,
                         particle_min_weight=MIN_W, nb_particles_per_cycle=500)),
    # source in the right-border rank, 7 particles per cycle (hundreds of
    # cycles, long exchange chains), kills at weight 1e-4
    "r4_c300_s7": (4, dict(nb_cells=300, nb_particles=3000,
                           x_ini="8.999999761581420898e-01",
                           particle_min_weight="9.999999747378751636e-05",
                           nb_particles_per_cycle=7)),
    # more particles than MAX_PARTICLES_VECTOR: lazy creation in chunks
    # (src/layer.cpp:89-121) while the loop runs; source in the middle rank
    "r3_c200_lazy": (3, dict(nb_cells=200, nb_particles=1200000,
                             x_ini="5.000000000000000000e-01",
                             particle_min_weight="9.999999939225290217e-09",
                             nb_particles_per_cycle=100000)),
    # one rank: no exchange at all
    "r1_c100": (1, dict(nb_cells=100, nb_particles=5000, x_ini=X_INI,
                        particle_min_weight=MIN_W, nb_particles_per_cycle=1000)),
}


def config_text(c):
    return ("x_min: 0.000000000000000000e+00\n"
            "x_max: 1.000000000000000000e+00\n"
            f"x_ini: {c['x_ini']}\n"
            f"particle_min_weight: {c['particle_min_weight']}\n"
            f"nb_cells: {c['nb_cells']}\n"
            f"nb_particles: {c['nb_particles']}\n"
            "cycle_time: 1.e-05\n"
            f"nb_particles_per_cycle: {c['nb_particles_per_cycle']}\n"
            "nthread: 1\n"
            "statistics_cycle_time: 1.e-1\n")


def stats_summary(text, ranks):
    lines = text.splitlines()
    rows = {r: [] for r in range(ranks)}
    for ln in lines[1:]:
        f = [v.strip() for v in ln.split(",")]
        rows[int(f[0])].append(int(f[-2]))
    return {"header": lines[0], "rows_per_rank": [len(rows[r]) for r in range(ranks)],
            "nb_cycles_per_rank": [sum(rows[r]) for r in range(ranks)]}


def main():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "refmain"], check=True)
    if not os.path.exists(EXE):
        sys.exit("oracle/_ref/mcmpi_main was not built (reference sources or MPI missing)")
    host = socket.gethostname()
    for name, (ranks, c) in CASES.items():
        d = os.path.join(OUT, name)
        os.makedirs(d, exist_ok=True)
        with tempfile.TemporaryDirectory() as t:
            cfg = os.path.join(t, "config.yaml")
            with open(cfg, "w") as f:
                f.write(config_text(c))
            r = subprocess.run([MPIRUN, "-n", str(ranks), EXE, "config.yaml", "sync"], cwd=t,
                               capture_output=True, text=True, timeout=1800)
            if r.returncode != 0:
                sys.exit(f"{name}: mpirun failed: {r.stderr[-2000:]}")
            shutil.copy(cfg, os.path.join(d, "config.yaml"))
            shutil.copy(os.path.join(t, "out", "weights.csv"), os.path.join(d, "weights.csv"))
            shutil.copy(os.path.join(t, "WA.out"), os.path.join(d, "WA.out"))
            oc = open(os.path.join(t, "out", "config.yaml")).read()
            oc = oc.replace(f"hostname: {host}\n", "hostname: HOSTNAME\n")
            with open(os.path.join(d, "out_config.yaml"), "w") as f:
                f.write(oc)
            s = stats_summary(open(os.path.join(t, "out", "stats.csv")).read(), ranks)
            s["ranks"] = ranks
            s["wall_s_printed"] = float(r.stdout.split()[0])
            with open(os.path.join(d, "summary.json"), "w") as f:
                json.dump(s, f, indent=1)
        print(name, ranks, s["nb_cycles_per_rank"], f"{s['wall_s_printed']:.3f} s", flush=True)


if __name__ == "__main__":
    main()
