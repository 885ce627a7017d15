"""The gsph command line (the reference's `gsph --benchmark N`, main.cpp:42-54):
its report keeps the reference's labels (sph_hybrid_system.cpp:397-429), and
its FP64 state checksum, printed at 17 significant digits, equals the C ABI's
for the same scene and frame count."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "fast-general-gpu-sph-framework_amd" / "_build" / "gsph"
LABELS = ("Total wall time:", "Average frame time:", "FPS:", "State checksum:", "Hybrid split:",
          "judgeTask isSame SMS pairs:", "arrange/grid :", "density      :", "force        :",
          "total kernel :")


def run(args, cwd):
    print(f"\n[gpu-case] gsph {' '.join(args)}", flush=True)
    r = subprocess.run([str(EXE)] + args, cwd=cwd, capture_output=True, text=True, timeout=200)
    assert r.returncode == 0, r.stderr
    for label in LABELS:
        assert label in r.stdout, label
    cs = re.search(r"State checksum: (.*)", r.stdout).group(1)
    return r.stdout, np.array([float(x) for x in cs.split(",")])


def test_cli_report_and_checksum(gsph, tmp_path):
    out, cs = run(["--config", "0", "--benchmark", "12"], tmp_path)
    p, pos = gsph.scene_config(0)
    s = gsph.System(p, pos)
    s.step(12)
    ref = s.checksum()
    s.close()
    assert np.array_equal(cs, ref), (cs, ref)
    assert re.search(r"Hybrid split: 32768 TRA particles, 0 SMS tasks", out)


def test_cli_no_lean_passes_same_checksum(tmp_path):
    """--no-lean-passes sends every tile to the general kernels: same state."""
    _, cs = run(["--config", "0", "--benchmark", "12"], tmp_path)
    _, cs_g = run(["--config", "0", "--benchmark", "12", "--no-lean-passes"], tmp_path)
    assert np.array_equal(cs, cs_g), (cs, cs_g)


def test_cli_default_scene_and_autotune(tmp_path):
    """No --scene, run outside the repository: the reference's own scene
    shipped with the package; --autotune changes the launch options and the
    split, never the state."""
    out, cs = run(["--benchmark", "6"], tmp_path)
    assert "Number of particles: 3944312" in out
    out_t, cs_t = run(["--autotune", "--benchmark", "6"], tmp_path)
    assert "Autotuned:" in out_t
    assert np.array_equal(cs, cs_t), (cs, cs_t)
