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


def _run(exe, work, nstp, nblk=1, nskip=0):
    shutil.copytree(os.path.join(DECKS, "n26g"), work, copy_function=shutil.copyfile)
    for t in ("up", "down"):
        shutil.copyfile(os.path.join(work, f"n26.{t}.x.res.0"), os.path.join(work, f"n26.{t}.x.0"))
    with open(os.path.join(work, "n26.in"), "w") as f:
        f.write(f"vmc {nblk} {nstp} 0.05 {nskip}\nend\n")
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


def _numbers(text):
    out = []
    for tok in text.split():
        try:
            out.append(float(tok))
        except ValueError:
            pass   # the labels of the rho_k files
    return np.array(out)


def _written(work):
    """What the run wrote besides the restart file (name -> values): the
    configuration stack and last walker (write_conf), the random seeds and
    the rho_k averages."""
    out = {}
    for fn in sorted(os.listdir(work)):
        if ".x." in fn or fn == "seed" or ".rhok_" in fn:
            with open(os.path.join(work, fn)) as f:
                out[fn] = _numbers(f.read())
    return out


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "refqmc_hipvmc")),
                    reason="device VMC driver not built (needs the reference sources)")
@pytest.mark.parametrize("nblk,nstp,nskip", [(3, 100, 0), (2, 90, 30)])
def test_device_vmc_driver_matches_reference_driver(tmp_path, nblk, nstp, nskip):
    """The reference driver with its vmc replaced by
    2degas_amd/fortran/vmc_hip.f, the Fortran caller of qmc_vmc_start and
    qmc_vmc_sweep (walker and random stream to the device, a sweep per
    block or per nskip steps, block sums into averages(2), the walker back
    into COMMON for putconf), against the reference driver on the CPU: the
    restart file (block averages, errors, RNG state), the configurations
    write_conf stores, the seeds and the rho_k averages agree."""
    print(f"\n[gpu] device vmc driver {nblk}x{nstp} nskip {nskip}", flush=True)
    cpu, _ = _run("refqmc", str(tmp_path / "cpu"), nstp, nblk, nskip)
    gpu, out = _run("refqmc_hipvmc", str(tmp_path / "gpu"), nstp, nblk, nskip)
    assert cpu.shape == gpu.shape
    rel = np.abs(cpu - gpu) / np.maximum(np.abs(cpu), 1e-300)
    print(f"   restart values {cpu.size} max rel diff {np.max(rel):.2e}", flush=True)
    assert np.max(rel) < 1e-8
    wc, wg = _written(str(tmp_path / "cpu")), _written(str(tmp_path / "gpu"))
    assert sorted(wc) == sorted(wg) and wc
    for fn in wc:
        assert wc[fn].shape == wg[fn].shape, fn
        d = np.max(np.abs(wc[fn] - wg[fn]) / np.maximum(np.abs(wc[fn]), 1.0))
        print(f"   {fn}: {wc[fn].size} values, max diff {d:.2e}", flush=True)
        assert d < 1e-8, fn
