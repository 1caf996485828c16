"""C++ host side (mcgpu::Layer over the C ABI): the compiled C++ tests under
tests/cpp, and the mcgpu_run driver."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "tests", "cpp", "_build")


def _run(args, cwd, timeout=600):
    out = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=timeout)
    print(out.stdout[-3000:], out.stderr[-3000:], flush=True)
    return out


def test_yaml_loader(tmp_path):
    exe = os.path.join(BUILD, "test_yaml")
    assert os.path.exists(exe), "build first (__graft_entry__.build())"
    out = _run([exe, str(tmp_path / "config.yaml")], str(tmp_path))
    assert out.returncode == 0 and "all ok" in out.stdout


def test_wire_particle_matches_reference_layout(tmp_path):
    """The 24-byte FP32 Particle of the reference (its MPI datatype's layout):
    the C++ WireParticle of the serialization literal of
    src/test_async_serialization.cpp:29-31 has the bytes numpy packs for the
    same record in the FP32 oracle's particle dtype, and FP32 -> FP64 -> FP32
    is the identity."""
    import numpy as np
    import pyoracle
    exe = os.path.join(BUILD, "test_wire")
    assert os.path.exists(exe), "build first (__graft_entry__.build())"
    out = _run([exe, str(tmp_path / "wire.bin")], str(tmp_path))
    assert out.returncode == 0 and "all ok" in out.stdout
    rec = np.zeros(1, pyoracle.P32)
    rec[0] = (5127801, 0.172634512365276, 123.12342351, -1231.123486123, 2134512)
    assert (tmp_path / "wire.bin").read_bytes() == rec.tobytes()


@pytest.mark.gpu
def test_layer_against_oracle(tmp_path):
    """mcgpu::Layer (GPU) vs the oracle's FP64 Layer: decompose_domain,
    lazy creation, simulate(n) sequences, escape buffers, nb_disabled,
    weights, WA.out -- for both kernels."""
    exe = os.path.join(BUILD, "test_layer_gpu")
    assert os.path.exists(exe)
    out = _run([exe], str(tmp_path))
    assert out.returncode == 0 and "all ok" in out.stdout


@pytest.mark.gpu
def test_mcgpu_run_cli_on_reference_config(tmp_path):
    """The reference's config.yaml source (1000 cells, default law; 2e4
    particles) through mcgpu_run: every particle is disabled, WA.out and
    out/{weights.csv,config.yaml,stats.csv} are written, and the weights
    equal the oracle's FP64 Layer run through the same cycles."""
    import numpy as np
    cfg = tmp_path / "config.yaml"
    cfg.write_text("x_min: 0.000000000000000000e+00\nx_max: 1.000000000000000000e+00\n"
                   "x_ini: 7.071067690849304199e-01\nparticle_min_weight: 9.999999960041972003e-13\n"
                   "nb_cells: 1000\nnb_particles: 20000\ncycle_time: 1.e-05\n"
                   "nb_particles_per_cycle: 5000\nnthread: 1\nstatistics_cycle_time: 1.e-1\n")
    out = _run([os.path.join(ROOT, "mc-mpi_amd", "mcgpu_run"), str(cfg)], str(tmp_path))
    assert out.returncode == 0
    assert "disabled 20000" in out.stderr
    wa = np.loadtxt(tmp_path / "WA.out")
    w = np.loadtxt(tmp_path / "out" / "weights.csv", delimiter=",", skiprows=1)
    assert (tmp_path / "out" / "config.yaml").exists() and (tmp_path / "out" / "stats.csv").exists()
    assert wa.shape == (1000, 2) and w.shape == (1000, 3)
    import pyoracle
    L = pyoracle.OracleLayer.decompose_domain(64, 0.0, 1.0, 7.071067690849304199e-01, 1, 0, 1000,
                                              20000, 9.999999960041972003e-13)
    L.set_tally_exact(True)
    while L.nb_disabled < 20000:
        L.simulate(5000, 8)
    np.testing.assert_array_equal(w[:, 2], L.weights_absorbed / L.dx)


@pytest.mark.gpu
def test_sync_cycle_against_oracle(tmp_path):
    """mcgpu::SyncCycle (every rank of a decomposition on one GPU) vs the
    oracle's synchronous worker loop: cycles, events, and per rank the
    particles, escape buffers, nb_disabled and weights, bit for bit."""
    exe = os.path.join(BUILD, "test_cycle_gpu")
    assert os.path.exists(exe)
    out = _run([exe], str(tmp_path))
    assert out.returncode == 0 and "all ok" in out.stdout


@pytest.mark.gpu
def test_mcgpu_run_five_ranks_on_one_gpu(tmp_path):
    """`mcgpu_run config.yaml --ranks 5` stands in for `mpirun -n 5 ./main
    config.yaml sync`: out/weights.csv holds the five ranks' tallies in rank
    order (proc column 0..4) and equals the oracle's five-rank loop."""
    import numpy as np
    import pyoracle
    cfg = tmp_path / "config.yaml"
    cfg.write_text("x_min: 0.000000000000000000e+00\nx_max: 1.000000000000000000e+00\n"
                   "x_ini: 7.071067690849304199e-01\nparticle_min_weight: 9.999999960041972003e-13\n"
                   "nb_cells: 1000\nnb_particles: 20000\ncycle_time: 1.e-05\n"
                   "nb_particles_per_cycle: 500\nnthread: 1\nstatistics_cycle_time: 1.e-1\n")
    out = _run([os.path.join(ROOT, "mc-mpi_amd", "mcgpu_run"), str(cfg), "--ranks", "5"],
               str(tmp_path))
    assert out.returncode == 0 and "disabled 20000" in out.stderr and "ranks 5" in out.stderr
    w = np.loadtxt(tmp_path / "out" / "weights.csv", delimiter=",", skiprows=1)
    assert w.shape == (1000, 3)
    assert np.array_equal(w[:, 0], np.repeat(np.arange(5), 200))
    layers, cycles = pyoracle.sync_run(64, 0.0, 1.0, 7.071067690849304199e-01, 5, 1000, 20000,
                                       9.999999960041972003e-13, 500, nthread=8, tally_exact=True)
    want = np.concatenate([L.weights_absorbed for L in layers]) / layers[0].dx
    np.testing.assert_array_equal(w[:, 2], want)
    _check_stats(tmp_path / "out" / "stats.csv", 5, cycles)


def _check_stats(path, ranks, cycles, period=0.1):
    """out/stats.csv of a run: the reference's header, then per rank its
    statistics windows (src/worker_sync.cpp:25-27,121-133): contiguous in
    time, every window but the last longer than the statistics period, the
    cycles counted adding up to all cycles but the one that ended the loop,
    and the phase times adding up to the window's length."""
    lines = path.read_text().splitlines()
    assert lines[0] == ("rank, starttime, endtime, time_comp, time_send, time_recv, "
                        "time_idle, nb_cycles, ")
    rows = {r: [] for r in range(ranks)}
    for ln in lines[1:]:
        assert ln.endswith(", ")
        f = [v.strip() for v in ln.split(",")[:-1]]
        assert len(f) == 8
        rows[int(f[0])].append([float(v) for v in f[1:7]] + [int(f[7])])
    for r in range(ranks):
        w = rows[r]
        assert w and sum(x[6] for x in w) == cycles - 1, (r, w)
        for a, b in zip(w, w[1:]):
            assert a[1] == b[0]                      # contiguous
            assert a[1] - a[0] > period              # closed after the period
        for x in w:
            assert x[1] >= x[0] and min(x[2:6]) >= 0
            assert abs(sum(x[2:6]) - (x[1] - x[0])) <= 1e-3 + 1e-2 * (x[1] - x[0])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["r1_c100", "r2_c1000", "r3_c200_lazy", "r4_c300_s7",
                                  "r5_c1000"])
def test_mcgpu_run_on_reference_fixture_configs(tmp_path, case):
    """`mcgpu_run config.yaml --ranks R` on the configs the reference's own
    program ran for tests/golden/sync/: out/config.yaml equals the
    reference's byte for byte (hostname aside); out/weights.csv equals the
    FP64 oracle's loop (the FP32 oracle's loop equals the reference program,
    tests/test_sync_reference.py), with the oracle's number of cycles;
    out/stats.csv has the reference's layout and window semantics."""
    import json
    import shutil
    import socket
    import numpy as np
    import pyoracle
    g = os.path.join(ROOT, "tests", "golden", "sync", case)
    summary = json.load(open(os.path.join(g, "summary.json")))
    R = summary["ranks"]
    shutil.copy(os.path.join(g, "config.yaml"), tmp_path / "config.yaml")
    out = _run([os.path.join(ROOT, "mc-mpi_amd", "mcgpu_run"), "config.yaml", "--ranks", str(R)],
               str(tmp_path))
    assert out.returncode == 0
    fields = dict(kv.split(" ", 1) for kv in
                  [ln for ln in out.stderr.splitlines() if ln.startswith("events ")][-1].split(", "))
    cfg = {k.strip(): v.strip() for k, v in
           (ln.split(":", 1) for ln in open(os.path.join(g, "config.yaml")))}
    got = (tmp_path / "out" / "config.yaml").read_text().replace(
        f"hostname: {socket.gethostname()}\n", "hostname: HOSTNAME\n")
    assert got == open(os.path.join(g, "out_config.yaml")).read()
    layers, cycles = pyoracle.sync_run(64, float(cfg["x_min"]), float(cfg["x_max"]),
                                       float(cfg["x_ini"]), R, int(cfg["nb_cells"]),
                                       int(cfg["nb_particles"]),
                                       float(cfg["particle_min_weight"]),
                                       int(cfg["nb_particles_per_cycle"]), nthread=8,
                                       tally_exact=True)
    assert int(fields["cycles"]) == cycles
    assert int(fields["disabled"]) == int(cfg["nb_particles"])
    w = np.loadtxt(tmp_path / "out" / "weights.csv", delimiter=",", skiprows=1)
    want = np.concatenate([L.weights_absorbed for L in layers]) / layers[0].dx
    np.testing.assert_array_equal(w[:, 2], want)
    _check_stats(tmp_path / "out" / "stats.csv", R, cycles)
