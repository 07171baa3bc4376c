"""bench.py without a GPU: its workloads are well-formed channel samplers of
the stated sizes, and it refuses more than one GPU."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


@pytest.fixture(scope="module")
def bench():
    import bench as b
    return b


def test_workloads_are_valid_samplers(bench):
    for name, w in bench.WORKLOADS.items():
        specs = bench.specs_of(name)
        ref = bench.mode_of(name) == bench.lbc.MODE_REF
        forcing = "forcing" in w
        assert len(specs) == (1 if ref else len(w["instances"])) >= 1, name
        assert w["storages"] and set(w["storages"]) <= {"f64", "f32", "f16",
                                                        "bf16"}
        for sp in specs:
            # the square keeps two fluid nodes from every boundary
            assert 2 <= sp.xmin < sp.xmax <= sp.Lx - 3, name
            assert 2 <= sp.ymin < sp.ymax <= sp.Ly - 3, name
            assert sp.tau > 0.5 and sp.n_clones >= 1
            if forcing:  # random-field force in the first forced_steps
                assert sp.beta > 0 and 0 < sp.forced_steps < sp.steps
                assert sp.field_eps > 0 and sp.n_snapshots == 0
            elif ref:  # body force, snapshot perturbation, no noise
                assert sp.beta > 0 and sp.n_snapshots > 0 and sp.eps > 0
            else:
                assert sp.noise_sigma > 0 and sp.beta == 0


def test_workload_sizes(bench):
    sizes = {name: [(sp.Lx, sp.Ly, sp.n_clones, sp.steps)
                    for sp in bench.specs_of(name)]
             for name in bench.WORKLOADS}
    assert sizes["channel400x100"] == [(400, 100, 4096, 200)]
    assert sizes["channel200x50"] == [(200, 50, 16, 100)]
    assert sizes["channel400x100-n512"] == [(400, 100, 512, 200)]
    assert sizes["channel800x200"] == [(800, 200, 2048, 400)]
    # the reference driver's box for L = 32: Dx = Dy = 129, dT*T0 = 320
    assert sizes["ref129"] == [(129, 129, 4096, 320)]
    assert sizes["forcing129"] == [(129, 129, 4096, 320)]
    assert bench.specs_of("forcing129")[0].forced_steps == 160
    (r,) = bench.specs_of("ref129")
    assert (r.xmin, r.xmax, r.ymin, r.ymax) == (64, 96, 48, 80)
    b8 = bench.specs_of("batched8")
    assert sum(sp.n_clones for sp in b8) == 4096
    assert sorted({sp.bias for sp in b8}) == [0.002, 0.01]
    for sp in b8:  # square side Ly/5 centred at (Lx/4, Ly/2)
        assert sp.xmax - sp.xmin == sp.Ly // 5
        assert (sp.xmin + sp.xmax) // 2 == sp.Lx // 4
        assert (sp.ymin + sp.ymax) // 2 == sp.Ly // 2
    capped = bench.specs_of("batched8", max_clones=256)
    assert [sp.n_clones for sp in capped] == [min(256, sp.n_clones) for sp in b8]
    assert bench.DEFAULT_WORKLOAD == "channel400x100"


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
