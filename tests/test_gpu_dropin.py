"""The drop-in: the reference's own Fortran driver (main + input +
sonaseppia + vmc, built from the reference sources into oracle/_ref)
linked with 2degas_amd/fortran/compute_properties_hip.f instead of its own
compute_properties, against the same driver on the CPU."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import DECKS, ROOT

pytestmark = pytest.mark.gpu
REF = os.path.join(ROOT, "oracle", "_ref")


def _run(exe, work, nstp):
    shutil.copytree(os.path.join(DECKS, "n26g"), work, copy_function=shutil.copyfile)
    for t in ("up", "down"):
        shutil.copyfile(os.path.join(work, f"n26.{t}.x.res.0"), os.path.join(work, f"n26.{t}.x.0"))
    with open(os.path.join(work, "n26.in"), "w") as f:
        f.write(f"vmc 1 {nstp} 0.05 0\nend\n")
    r = subprocess.run([os.path.join(REF, exe)], cwd=work, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=300)
    assert r.returncode == 0, r.stdout.decode()[-2000:]
    with open(os.path.join(work, "n26.res")) as f:
        vals = np.array([float(v) for v in f.read().replace("vmc", " ").split()])
    return vals, r.stdout.decode()


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "refqmc_hip")),
                    reason="drop-in driver not built (needs the reference sources)")
def test_reference_vmc_driver_with_gpu_compute_properties(tmp_path):
    """300 VMC steps of the reference driver: the block averages, the RNG
    state written to the restart file and the acceptance sum agree."""
    print("\n[gpu] drop-in reference driver", flush=True)
    cpu, _ = _run("refqmc", str(tmp_path / "cpu"), 300)
    gpu, out = _run("refqmc_hip", str(tmp_path / "gpu"), 300)
    assert cpu.shape == gpu.shape
    rel = np.abs(cpu - gpu) / np.maximum(np.abs(cpu), 1e-300)
    print(f"   restart values {cpu.size} max rel diff {np.max(rel):.2e}", flush=True)
    assert np.max(rel) < 1e-8
