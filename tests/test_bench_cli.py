"""bench.py's command-line contract and host logic, checked without a GPU:
more than one GPU is refused with a message, and the batched implicit
workload integrates every problem to steps*dt for every method, also after
the shared network has been trained in the same run."""
import importlib.util
import os
import subprocess
import sys
import types

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = run("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout.strip() == ""


class _Stub:
    """Host-only stand-in for the package: records the end time each
    integrator is asked to reach and the training steps requested."""

    MMS_DIRICHLET = 1

    def __init__(self):
        self.evolve_targets = []
        self.train_steps = []
        self.train_ratios = []
        stub = self

        class Context:
            def close(self):
                pass

        class Vec:
            def __init__(self, n, value=0.0):
                self.n, self.value = n, value

            def get(self):
                return np.full(self.n, self.value)

        class Problem:
            def __init__(self, ctx, kind, nx, ny, kx, ky):
                self.n = nx * ny

            def initial(self):
                return Vec(self.n)

            def exact(self, t):
                return Vec(self.n, 1.0)

            def destroy(self):
                pass

        class LearnedPrec:
            def __init__(self, ctx, dil, symmetric):
                pass

            def init_identity(self, *a):
                pass

            def set_output_scale(self, s):
                pass

            def train(self, nsamp, tn, ratio, R, E, steps, batch, lr, seed, ratio_y=None):
                stub.train_steps.append(steps)
                return np.ones(max(steps, 50))

            def params(self):
                return np.zeros(4, np.float32)

            def set_params(self, p):
                pass

            def attach(self, G):
                pass

        class Integrator:
            def __init__(self, G, method, y0, ls, maxl, precond):
                self.LS = None
                self.y = Vec(G.n)

            def __getattr__(self, name):
                return lambda *a: 0

            def evolve(self, tout):
                stub.evolve_targets.append(tout)
                return 0, tout

            def stats(self):
                return {"nli": 1}

            def fallbacks(self):
                return 0

        self.Context, self.Problem, self.LearnedPrec, self.Integrator = Context, Problem, LearnedPrec, Integrator
        def training_set(ctx, nsamp, tn, ratio, reltol, seed, ratio_y=None):
            stub.train_ratios.append((ratio, ratio if ratio_y is None else ratio_y))
            return None, None, 0, 0

        self.training_set = training_set
        self.lib = lambda: types.SimpleNamespace(SUNLinSol_SPFGMRSetMaxRestartDim=lambda *a: 0)


def test_c5_problems_reach_steps_times_dt_after_training(capsys):
    """The shared network's training loop once rebound the variable that held
    the DIRK step count, so every problem was integrated to the training-step
    count times dt (333x too far).  Each method must integrate to c5_steps*dt."""
    bench = load_bench()
    bench.hip_sync = lambda: None
    stub = _Stub()
    args = bench.parse_args(["--workload", "c5", "--c5-n", "3", "--c5-workers", "1", "--c5-steps", "3",
                             "--c5-methods", "jacobi,learned-sym,mg", "--nsamp", "4", "--c5-classes", "0",
                             "--sym-train-steps", "1000"])
    bench.c5_workload(stub, args)
    assert stub.train_steps == [1000]
    # jacobi, learned-sym, mg, then jacobi on one worker is skipped (workers == 1): 3 methods x 3 problems
    assert len(stub.evolve_targets) == 9
    assert all(t == 3 * 1e-3 for t in stub.evolve_targets), stub.evolve_targets
    out = capsys.readouterr().out
    assert "3 steps" in out and "1000 steps" not in out


def test_c5_classes_train_one_network_per_anisotropy_class():
    """--c5-classes K trains K networks on the anisotropic operators of K
    equal-count classes of r_x/r_y, ordered by anisotropy; without it one
    isotropic network."""
    import numpy as np
    bench = load_bench()
    bench.hip_sync = lambda: None
    probs = bench.c5_problems(6, 20260102)
    for classes, want in ((0, 1), (3, 3)):
        stub = _Stub()
        args = bench.parse_args(["--workload", "c5", "--c5-n", "6", "--c5-workers", "1", "--c5-steps", "1",
                                 "--c5-methods", "learned-sym,learned-sym-pcg", "--nsamp", "4",
                                 "--c5-classes", str(classes), "--c5-mag-classes", "1"])
        bench.c5_workload(stub, args)
        assert len(stub.train_ratios) == want  # the PCG method shares the FGMRES method's networks
        assert len(stub.evolve_targets) == 12
        if classes == 0:
            assert stub.train_ratios[0][0] == stub.train_ratios[0][1]
            continue
        aniso = sorted(p["kx"] / p["ky"] for p in probs)  # square grids: r_x/r_y = kx/ky
        got = [rx / ry for rx, ry in stub.train_ratios]
        for c in range(3):
            assert np.isclose(got[c], np.exp(np.mean(np.log(aniso[2 * c:2 * c + 2])))), (c, got, aniso)


def test_c5_nets_file_is_written_then_reused(tmp_path):
    """--c5-nets: the first run trains and writes the networks, a second run
    with the same training operators loads them instead of training; other
    operators train again."""
    bench = load_bench()
    bench.hip_sync = lambda: None
    path = str(tmp_path / "nets.npz")
    argv = ["--workload", "c5", "--c5-n", "4", "--c5-workers", "1", "--c5-steps", "1",
            "--c5-methods", "learned-sym-pcg", "--nsamp", "4", "--c5-nets", path, "--c5-mag-classes", "1"]
    counts = []
    for extra in (["--c5-classes", "0"], ["--c5-classes", "0"], ["--c5-classes", "2"]):
        stub = _Stub()
        bench.c5_workload(stub, bench.parse_args(argv + extra))
        counts.append(len(stub.train_ratios))
    assert counts == [1, 0, 2]


def test_c5_magnitude_classes_split_each_anisotropy_class():
    """--c5-mag-classes M splits each of the K anisotropy classes into M
    classes by r_x r_y: K*M networks, each trained at its class's
    geometric-mean operator."""
    import numpy as np
    bench = load_bench()
    bench.hip_sync = lambda: None
    stub = _Stub()
    args = bench.parse_args(["--workload", "c5", "--c5-n", "8", "--c5-workers", "1", "--c5-steps", "1",
                             "--c5-methods", "learned-sym-pcg", "--nsamp", "4", "--c5-classes", "2",
                             "--c5-mag-classes", "2"])
    bench.c5_workload(stub, args)
    assert len(stub.train_ratios) == 4
    probs = bench.c5_problems(8, 20260102)
    g = 1767732205903.0 / 4055673282236.0 * 1e-3
    rxy = [(g * p["kx"] * (p["nx"] + 1) ** 2, g * p["ky"] * (p["ny"] + 1) ** 2) for p in probs]
    order = sorted(range(8), key=lambda k: rxy[k][0] / rxy[k][1])
    for c, half in enumerate((order[:4], order[4:])):
        half = sorted(half, key=lambda k: rxy[k][0] * rxy[k][1])
        for m, grp in enumerate((half[:2], half[2:])):
            want = [np.exp(np.mean([np.log(rxy[k][a]) for k in grp])) for a in (0, 1)]
            assert np.allclose(stub.train_ratios[2 * c + m], want, rtol=1e-12)


def test_c5_defaults_are_twelve_class_networks_and_every_method():
    bench = load_bench()
    args = bench.parse_args(["--workload", "c5"])
    assert (args.c5_classes, args.c5_mag_classes) == (4, 3)
    assert set(args.c5_methods.split(",")) == {"jacobi", "mg", "learned-sym-pcg", "learned-sym"}
