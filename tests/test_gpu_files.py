"""The launcher's file flow on the GPU: inputs read from xmcml files, runs in
backup portions (LaunchOMPWithBackups + DoBackup,
base/xmcml/xmcmlLauncher/launcher_mpi.cpp:311-380) and resumes from a backup
(main.cpp:106-127).  Photon p always draws stream block p, so a run split
into portions, or resumed from any backup, adds the same photons as one
uninterrupted launch: detector and time-bin counts are identical, weights
agree to FP64 summation-order rounding (1e-9 relative)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from xmcml_amd import engine, files, scenes
from xmcml_amd.abi import InputHandle, OutputHandle, Rng

pytestmark = pytest.mark.gpu

N = 40_000
PORTION = 10_000


@pytest.fixture(scope="module")
def slab_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("slab")
    spec = scenes.slab_scene(N)
    files.write_input_xml(spec, d / "slab.xml", photons=N)
    files.write_surface_file(spec, d / "slab.surface")
    return d


def _close(a: OutputHandle, b: OutputHandle):
    assert np.array_equal(a.det_count(), b.det_count())
    assert np.array_equal(a.bin_count(), b.bin_count())
    np.testing.assert_allclose(a.det_weight(), b.det_weight(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(a.bin_weight(), b.bin_weight(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(a.absorption, b.absorption, rtol=1e-9, atol=1e-12)


def _one_launch(doc, spec):
    out = OutputHandle(spec, absorption=True)
    rng = Rng(777, 0)
    engine._check(engine.lib().xmc_launch(doc.ptr, out.ptr, C.byref(rng), None), "xmc_launch")
    assert rng.nextPhoton == N
    return out


def test_portions_and_resume_equal_one_launch(engine, slab_files, monkeypatch):
    monkeypatch.chdir(slab_files)
    doc = files.Document(slab_files / "slab.xml", slab_files / "slab.surface")
    spec = doc.spec()
    assert spec.termination == 0 and doc.input.numberOfPhotons == N
    whole = _one_launch(doc, spec)
    print(f"files: slab from xml, {N} photons in portions of {PORTION}", flush=True)

    parts = OutputHandle(spec, absorption=True)
    rng = Rng(777, 0)
    files.launch_with_backups(doc, parts, rng, PORTION)  # backups under the reference's names
    assert (rng.nextPhoton, parts.c.numberOfPhotons) == (N, N)
    _close(parts, whole)
    names = [files.backup_file_name(k * PORTION) for k in (1, 2, 3)]
    assert all((slab_files / n).exists() for n in names)

    # resume from the second backup (20,000 photons done)
    res = OutputHandle(spec, absorption=True)
    r2 = files.read_backup_file(slab_files / names[1], res)
    assert (r2.seed, r2.nextPhoton, res.c.numberOfPhotons) == (777, 2 * PORTION, 2 * PORTION)
    doc.input.numberOfPhotons = N - res.c.numberOfPhotons
    files.launch_with_backups(doc, res, r2, PORTION)
    assert (r2.nextPhoton, res.c.numberOfPhotons) == (N, N)
    _close(res, whole)


def test_run_files_writes_the_result(engine, slab_files):
    out = files.run_files(slab_files / "slab.xml", slab_files / "slab.surface", slab_files / "result.mcml.out",
                          seed=777, portion=15_000, backup_path=slab_files / "run.bk")
    doc = files.Document(slab_files / "slab.xml", slab_files / "slab.surface")
    spec = doc.spec()
    back = OutputHandle(spec, absorption=True, scattering=True, trajectory=True)
    files.read_backup_file(slab_files / "result.mcml.out", back, with_rng=False)
    assert back.c.numberOfPhotons == N
    assert np.array_equal(back.absorption, out.absorption) and np.array_equal(back.scattering, out.scattering)
    assert np.array_equal(back.det_count(), out.det_count())
    assert out.scattering.sum() > 0  # the result carries the scattering grid, as the reference's does
    # and every detector's trajectory map (InitOutput allocates them, launcher_omp.cpp:33-34; the writer
    # stores them, writer.cpp:431-471): equal to the float oracle's maps of the same photons
    assert np.array_equal(back.trajectory, out.trajectory) and out.trajectory.sum() > 0
    o = OutputHandle(spec, absorption=True, scattering=True, trajectory=True)
    ol.run_oracle(spec, InputHandle(spec), o, 777, 0, N, 32, 4, scattering=True)  # 4 thread copies of 160 MB maps
    assert np.array_equal(out.trajectory, o.trajectory)
    assert np.array_equal(out.det_count(), o.det_count())
    whole = _one_launch(doc, spec)
    _close(out, whole)


def test_run_files_with_biasing_takes_or_writes_the_table_file(engine, tmp_path, monkeypatch):
    """GetWeightIntegralTable (weight_integral.cpp:16-31): without a table file
    the tables are computed on the GPU and written under the reference's name
    in the working directory; a run given that file (the launcher's -t) reads
    it back and adds the same photons."""
    monkeypatch.chdir(tmp_path)
    spec = scenes.biased(scenes.slab_scene(N), target=(2.0, 0.0, 0.0), factor=0.5, integral_precision=16,
                         table_precision=1024)
    files.write_input_xml(spec, tmp_path / "b.xml", photons=N)
    files.write_surface_file(spec, tmp_path / "b.surface")
    print("files: biased slab, tables computed and written", flush=True)
    a = files.run_files(tmp_path / "b.xml", tmp_path / "b.surface", tmp_path / "a.mcml.out", seed=11, portion=N)
    table = tmp_path / files.WEIGHT_TABLE_FILE
    assert table.exists()
    doc = files.Document(tmp_path / "b.xml", tmp_path / "b.surface")
    doc.read_weight_table_file(table)
    for g, el in doc.spec().weight_tables:
        assert np.array_equal(el, engine.weight_table(g, 0.5, 16, 1024))
    table.rename(tmp_path / "given.bin")
    print("files: biased slab, tables from the file", flush=True)
    b = files.run_files(tmp_path / "b.xml", tmp_path / "b.surface", tmp_path / "b.mcml.out", seed=11, portion=N,
                        weight_table_path=tmp_path / "given.bin")
    assert not table.exists()
    _close(a, b)
    assert a.det_weight().sum() > 0
