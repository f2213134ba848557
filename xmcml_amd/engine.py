"""ctypes binding of the engine library (libxmcml_amd.so, include/xmcml.h).

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.  ``Scene`` mirrors the reference's split between building the
geometry once (GenerateBVHTree, base/xmcml/xmcml/mcml_intersection.cpp:316)
and running photon portions into accumulated outputs (LaunchOMP,
base/xmcml/xmcmlLauncher/launcher_omp.cpp:123).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Optional

import numpy as np

from .abi import (TRACE_DTYPE, InputHandle, Options, OutputHandle, Rng, SceneSpec, Stats)

# XMC_LIBRARY selects another build of the library (the sanitized one of the CPU test runs)
LIB_PATH = Path(os.environ.get("XMC_LIBRARY") or Path(__file__).resolve().parent / "libxmcml_amd.so")

_P = C.c_void_p
_I = C.c_int
_U64 = C.c_uint64
_PD = C.POINTER(C.c_double)
_PI = C.POINTER(C.c_int)

_lib = None

EXPORTS = [
    "xmc_scene_create", "xmc_scene_destroy", "xmc_scene_run", "xmc_scene_sync", "xmc_scene_collect", "xmc_scene_stats",
    "xmc_scene_trace", "xmc_scene_query", "xmc_scene_query_device", "xmc_scene_bvh", "xmc_scene_set_launch", "xmc_scene_get_launch",
    "xmc_batch_create", "xmc_batch_destroy", "xmc_batch_run", "xmc_batch_collect", "xmc_batch_stats",
    "xmc_batch_set_launch", "xmc_batch_get_launch", "xmc_distance_grid", "xmc_slab_grid", "xmc_launch", "xmc_init_output",
    "xmc_free_output", "xmc_specular_reflectance", "xmc_device_name", "xmc_last_error", "xmc_version",
    # surrogate model (include/xmcml_surrogate.h)
    "xmc_mlp_create", "xmc_mlp_destroy", "xmc_mlp_param_count", "xmc_mlp_get_params", "xmc_mlp_set_params",
    "xmc_mlp_forward", "xmc_mlp_forward_layers", "xmc_mlp_backward", "xmc_mlp_get_grads", "xmc_mlp_adam", "xmc_mlp_train_step",
    "xmc_mlp_gemm_check", "xmc_mlp_train_sampled", "xmc_surrogate_encode", "xmc_mlp_answer", "xmc_dev_alloc", "xmc_dev_free", "xmc_dev_copy", "xmc_dev_sync",
    "xmc_weight_integrals", "xmc_weight_table_compute",
    # file formats and backups (include/xmcml_io.h)
    "xmc_read_input_file", "xmc_read_surface_file", "xmc_document_input", "xmc_document_biasing",
    "xmc_document_free", "xmc_write_surface_file", "xmc_write_output_file", "xmc_write_backup_file",
    "xmc_backup_file_name", "xmc_read_backup_file", "xmc_launch_with_backups", "xmc_write_weight_table_file",
    "xmc_read_weight_table_file",
]


class EngineError(RuntimeError):
    pass


def lib():
    """Load the engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise EngineError(f"{LIB_PATH} is missing: build it with `python -m xmcml_amd.build`")
    L = C.CDLL(str(LIB_PATH))

    def sig(name, res, args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args

    sig("xmc_scene_create", _I, [_P, _P, C.POINTER(_P)])
    sig("xmc_scene_destroy", _I, [_P])
    sig("xmc_scene_run", _I, [_P, _U64, _U64, _U64, _P])
    sig("xmc_scene_sync", _I, [_P])
    sig("xmc_scene_collect", _I, [_P, _P])
    sig("xmc_scene_stats", _I, [_P, _P])
    sig("xmc_scene_trace", _I, [_P, _U64, _U64, _U64, _P, _I, _PI])
    sig("xmc_scene_query", _I, [_P, _PD, _PD, _PD, _I, _I, _PI, _PD, _PI, _PI, _PD])
    sig("xmc_scene_query_device", _I, [_P, _P, _P, _P, _I, _I, _P, _P, _P, _P, _P, _P])
    sig("xmc_device_name", _I, [_I, C.c_char_p, _I, _PI])
    sig("xmc_scene_bvh", _I, [_P, _PI, _PI, _PI, _PD, _PI, _PI, _PI])
    sig("xmc_scene_set_launch", _I, [_P, _I, _I])
    sig("xmc_scene_get_launch", _I, [_P, _PI, _PI])
    sig("xmc_batch_create", _I, [_P, _I, _P, C.POINTER(_P)])
    sig("xmc_batch_destroy", _I, [_P])
    sig("xmc_batch_run", _I, [_P, _U64, _P])
    sig("xmc_batch_collect", _I, [_P, _P])
    sig("xmc_batch_stats", _I, [_P, _P])
    sig("xmc_batch_set_launch", _I, [_P, _I, _I, _I])
    sig("xmc_batch_get_launch", _I, [_P, _PI, _PI])
    sig("xmc_distance_grid", _I, [_P, _U64, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                  _PI, _P])
    sig("xmc_slab_grid", _I, [_P, _U64, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                              _PI, _P])
    sig("xmc_launch", _I, [_P, _P, _P, _P])
    sig("xmc_init_output", _I, [_P, _P])
    sig("xmc_free_output", None, [_P])
    sig("xmc_specular_reflectance", C.c_double, [_P])
    sig("xmc_last_error", C.c_char_p, [])
    sig("xmc_version", C.c_char_p, [])
    _F = C.c_float
    _PF = C.POINTER(C.c_float)
    sig("xmc_mlp_create", _I, [_P, _U64, C.POINTER(_P)])
    sig("xmc_mlp_destroy", _I, [_P])
    sig("xmc_mlp_param_count", _I, [_P, C.POINTER(C.c_int64)])
    sig("xmc_mlp_get_params", _I, [_P, _PF, C.c_int64])
    sig("xmc_mlp_set_params", _I, [_P, _PF, C.c_int64])
    sig("xmc_mlp_forward", _I, [_P, _P, _I, _P, _P])
    sig("xmc_mlp_forward_layers", _I, [_P, _P, _I, _P, _P])
    sig("xmc_mlp_backward", _I, [_P, _P, _P, _P, _I, _F, _PF, _P])
    sig("xmc_mlp_get_grads", _I, [_P, _PF, C.c_int64])
    sig("xmc_mlp_adam", _I, [_P, _F, _F, _F, _F, _F, _P])
    sig("xmc_mlp_train_step", _I, [_P, _P, _P, _P, _I, _F, _F, _F, _F, _F, _F, _P])
    sig("xmc_mlp_gemm_check", _I, [_PF, _PF, _I, _I, _I, _PF, _I, _I])
    sig("xmc_mlp_train_sampled", _I, [_P, _P, _P, _P, C.c_int64, _I, _U64, C.c_int64, _F, _F, _F, _F, _F, _F,
                                      _PF, _P])
    sig("xmc_surrogate_encode", _I, [_P, _P, C.c_int64, _P, _P, _P, _I, _P])
    sig("xmc_mlp_answer", _I, [_P, _P, C.c_int64, _PF, _PF, _I, _F, _P, _P, _P])
    sig("xmc_dev_alloc", _I, [_I, _U64, C.POINTER(_P)])
    sig("xmc_dev_free", _I, [_P])
    sig("xmc_dev_copy", _I, [_P, _P, _U64, _I])
    sig("xmc_dev_sync", _I, [_I])
    sig("xmc_weight_integrals", _I, [_PD, _I, C.c_double, C.c_double, _I, _PD, _I])
    sig("xmc_weight_table_compute", _I, [C.c_double, C.c_double, _I, _I, _PD, _I])
    _S = C.c_char_p
    sig("xmc_read_input_file", _I, [_S, C.POINTER(_P)])
    sig("xmc_read_surface_file", _I, [_P, _S])
    sig("xmc_document_input", _P, [_P])
    sig("xmc_document_biasing", _I, [_P, _P])
    sig("xmc_document_free", None, [_P])
    sig("xmc_write_surface_file", _I, [_P, _I, _S])
    sig("xmc_write_output_file", _I, [_P, _P, _S])
    sig("xmc_write_backup_file", _I, [_P, _P, _P, _S])
    sig("xmc_backup_file_name", _I, [_U64, C.c_char_p, _I])
    sig("xmc_read_backup_file", _I, [_S, _P, _P])
    sig("xmc_launch_with_backups", _I, [_P, _P, _P, _P, _U64, _S])
    sig("xmc_write_weight_table_file", _I, [_P, _S])
    sig("xmc_read_weight_table_file", _I, [_P, _S])
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != 0:
        msg = lib().xmc_last_error().decode(errors="replace")
        raise EngineError(f"{what} failed ({rc}): {msg}")


def _dp(a):
    return a.ctypes.data_as(_PD)


def _ip(a):
    return a.ctypes.data_as(_PI)


class Scene:
    """Geometry, optics and device tallies of one input, resident on the GPU."""

    def __init__(self, spec: SceneSpec, block_size: int = 0, blocks_per_cu: int = 0,
                 absorption: bool = True, device: int = 0, distance_grid: bool = True,
                 scattering: bool = False, trajectories: bool = False, trajectory_capacity: int = 0,
                 entry_grid: bool = True, slab_grid: bool = True, lattice_rule: bool = True):
        self.spec = spec
        self.input = InputHandle(spec)
        opt = Options(device, block_size, blocks_per_cu, 0 if absorption else 1, 0 if distance_grid else 1,
                      int(scattering), int(trajectories), trajectory_capacity, 0 if entry_grid else 1,
                      0 if slab_grid else 1, 0 if lattice_rule else 1)
        self.scattering, self.trajectories = scattering, trajectories
        h = _P()
        _check(lib().xmc_scene_create(self.input.ptr, C.byref(opt), C.byref(h)), "xmc_scene_create")
        self._h = h
        self.absorption = absorption

    def close(self):
        if getattr(self, "_h", None):
            lib().xmc_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, seed: int, first: int, n: int, stream: Optional[int] = None):
        _check(lib().xmc_scene_run(self._h, seed, first, n, stream), "xmc_scene_run")

    def sync(self):
        _check(lib().xmc_scene_sync(self._h), "xmc_scene_sync")

    def collect(self, out: Optional[OutputHandle] = None) -> OutputHandle:
        if out is None:
            out = OutputHandle(self.spec, absorption=self.absorption, scattering=self.scattering,
                               trajectory=self.trajectories)
        _check(lib().xmc_scene_collect(self._h, out.ptr), "xmc_scene_collect")
        return out

    def stats(self) -> dict:
        s = Stats()
        _check(lib().xmc_scene_stats(self._h, C.byref(s)), "xmc_scene_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def trace(self, seed: int, first: int, n: int, max_events: int = 4096):
        ev = np.zeros(n * max_events, dtype=TRACE_DTYPE)
        cnt = np.zeros(n, dtype=np.int32)
        _check(lib().xmc_scene_trace(self._h, seed, first, n, ev.ctypes.data, max_events, _ip(cnt)),
               "xmc_scene_trace")
        ev = ev.reshape(n, max_events)
        return [ev[i, :min(cnt[i], max_events)].copy() for i in range(n)], cnt

    def query(self, origins, directions, steps, brute: bool = False):
        o = np.ascontiguousarray(origins, dtype=np.float64)
        d = np.ascontiguousarray(directions, dtype=np.float64)
        s = np.ascontiguousarray(steps, dtype=np.float64)
        n = s.shape[0]
        found = np.zeros(n, np.int32)
        dist = np.zeros(n)
        surf = np.zeros(n, np.int32)
        tri = np.zeros(n, np.int32)
        nrm = np.zeros((n, 3))
        _check(lib().xmc_scene_query(self._h, _dp(o), _dp(d), _dp(s), n, int(brute), _ip(found), _dp(dist),
                                     _ip(surf), _ip(tri), _dp(nrm)), "xmc_scene_query")
        return dict(found=found, distance=dist, surface=surf, triangle=tri, normal=nrm)

    def set_launch(self, block: int, blocks_per_cu: int = 0):
        _check(lib().xmc_scene_set_launch(self._h, block, blocks_per_cu), "xmc_scene_set_launch")

    def get_launch(self):
        b, n = C.c_int(), C.c_int()
        _check(lib().xmc_scene_get_launch(self._h, C.byref(b), C.byref(n)), "xmc_scene_get_launch")
        return b.value, n.value

    def bvh(self) -> dict:
        nn, root, nt = C.c_int(), C.c_int(), C.c_int()
        _check(lib().xmc_scene_bvh(self._h, C.byref(nn), C.byref(root), C.byref(nt), None, None, None, None),
               "xmc_scene_bvh")
        boxes = np.zeros((nn.value, 6))
        links = np.zeros((nn.value, 4), np.int32)
        ts = np.zeros(nt.value, np.int32)
        ti = np.zeros(nt.value, np.int32)
        _check(lib().xmc_scene_bvh(self._h, C.byref(nn), C.byref(root), C.byref(nt), _dp(boxes), _ip(links),
                                   _ip(ts), _ip(ti)), "xmc_scene_bvh")
        return dict(root=root.value, boxes=boxes, links=links, tsurf=ts, tidx=ti)


class DeviceArray:
    """An array in device memory (float32 unless dtype says otherwise),
    allocated by the engine library (xmc_dev_alloc); upload() / numpy() copy
    through host memory."""

    def __init__(self, shape, data=None, device: int = 0, dtype=np.float32):
        self.shape = tuple(int(v) for v in shape)
        self.device = device
        self.dtype = np.dtype(dtype)
        self.nbytes = self.dtype.itemsize * int(np.prod(self.shape)) if self.shape else self.dtype.itemsize
        p = _P()
        _check(lib().xmc_dev_alloc(device, max(self.nbytes, 1), C.byref(p)), "xmc_dev_alloc")
        self._p = p
        if data is not None:
            self.upload(data)

    @property
    def ptr(self) -> _P:
        return self._p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        if self.nbytes:
            _check(lib().xmc_dev_copy(self._p, a.ctypes.data, self.nbytes, 1), "xmc_dev_copy")

    def numpy(self) -> np.ndarray:
        a = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _check(lib().xmc_dev_copy(a.ctypes.data, self._p, self.nbytes, 2), "xmc_dev_copy")
        return a

    def close(self):
        if getattr(self, "_p", None):
            lib().xmc_dev_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_name(device: int = 0) -> str:
    """The HIP device's name and architecture (xmc_device_name)."""
    buf = C.create_string_buffer(256)
    cus = C.c_int()
    _check(lib().xmc_device_name(device, buf, 256, C.byref(cus)), "xmc_device_name")
    return f"{buf.value.decode(errors='replace')} {cus.value}CU"


def device_sync(device: int = 0):
    _check(lib().xmc_dev_sync(device), "xmc_dev_sync")


def distance_grid(spec: SceneSpec, target_cells: int = 0) -> dict:
    """The conservative distance grid a scene builds (host computation)."""
    inp = InputHandle(spec)
    lo = (C.c_float * 3)()
    h, q = C.c_float(), C.c_float()
    dims = (C.c_int * 3)()
    _check(lib().xmc_distance_grid(inp.ptr, target_cells, lo, C.byref(h), C.byref(q), dims, None),
           "xmc_distance_grid")
    grid = np.zeros(dims[0] * dims[1] * dims[2], np.uint8)
    _check(lib().xmc_distance_grid(inp.ptr, target_cells, lo, C.byref(h), C.byref(q), dims, grid.ctypes.data),
           "xmc_distance_grid")
    return dict(lo=np.array(lo[:], np.float32), h=h.value, step_q=q.value, dims=tuple(dims[:]),
                q=grid.reshape(dims[2], dims[1], dims[0]))


def slab_grid(spec: SceneSpec, mean_step: float, target_cells: int = 0) -> dict:
    """The separating slabs a scene builds near its mesh (host computation):
    per cell the packed 64-bit word of xmc_slab_grid."""
    inp = InputHandle(spec)
    lo = (C.c_float * 3)()
    h, unit = C.c_float(), C.c_float()
    dims = (C.c_int * 3)()
    _check(lib().xmc_slab_grid(inp.ptr, target_cells, mean_step, lo, C.byref(h), C.byref(unit), dims, None),
           "xmc_slab_grid")
    words = np.zeros(dims[0] * dims[1] * dims[2], np.uint64)
    _check(lib().xmc_slab_grid(inp.ptr, target_cells, mean_step, lo, C.byref(h), C.byref(unit), dims,
                               words.ctypes.data), "xmc_slab_grid")
    return dict(lo=np.array(lo[:], np.float32), h=h.value, unit=unit.value, dims=tuple(dims[:]),
                words=words.reshape(dims[2], dims[1], dims[0]))


def weight_integrals(dots, anisotropy: float, attractive_factor: float, precision: int, device: int = 0):
    """ComputeWeightIntegral at each dot, on the GPU (FP64)."""
    d = np.ascontiguousarray(dots, dtype=np.float64)
    out = np.zeros(d.size)
    _check(lib().xmc_weight_integrals(_dp(d), d.size, anisotropy, attractive_factor, precision, _dp(out), device),
           "xmc_weight_integrals")
    return out


def weight_table(anisotropy: float, attractive_factor: float = 1.0, integral_precision: int = 128,
                 table_precision: int = 262144, device: int = 0):
    """The biasing's weight-integral table of one anisotropy (tablePrecision + 1 entries)."""
    out = np.zeros(table_precision + 1)
    _check(lib().xmc_weight_table_compute(anisotropy, attractive_factor, integral_precision, table_precision,
                                          _dp(out), device), "xmc_weight_table_compute")
    return out


def launch(spec: SceneSpec, out: OutputHandle, seed: int, first: int, n: int, absorption: bool = True) -> int:
    """The LaunchOMP replacement: one call, tallies added into ``out``.
    Returns the next photon index of the stream."""
    inp = InputHandle(spec, photons=n)
    rng = Rng(seed, first)
    opt = Options(0, 0, 0, 0 if absorption else 1, 0, 0, 0, 0, 0)
    _check(lib().xmc_launch(inp.ptr, out.ptr, C.byref(rng), C.byref(opt)), "xmc_launch")
    return rng.nextPhoton


class Batch:
    """Many configurations run by one persistent launch (replayed as a HIP
    graph); configuration c's budget is specs[c].photons."""

    def __init__(self, specs, block_size: int = 0, blocks_per_cu: int = 0, absorption: bool = False,
                 use_graph: bool = True, device: int = 0, tune: bool = False, tune_scale: float = 0.01):
        """tune: choose the launch shape with the autotuner (xmcml_amd/autotune.py) on a copy of
        the workload whose photon budgets are scaled by tune_scale, and apply it here; the
        choice and every candidate's kernel time are in self.tuning."""
        from .abi import Input
        self.specs = list(specs)
        self.device = device
        self.tuning = None
        self.inputs = [InputHandle(s) for s in self.specs]
        arr = (Input * len(self.inputs))(*[h.c for h in self.inputs])
        self._arr = arr
        opt = Options(device, block_size, blocks_per_cu, 0 if absorption else 1, 0, 0, 0, 0, 0)
        h = _P()
        _check(lib().xmc_batch_create(arr, len(self.inputs), C.byref(opt), C.byref(h)), "xmc_batch_create")
        self._h = h
        self.absorption = absorption
        if not use_graph or block_size or blocks_per_cu:
            self.set_launch(block_size, blocks_per_cu, use_graph)
        if tune:
            from . import autotune
            import dataclasses
            scaled = [dataclasses.replace(s, photons=max(1000, int(s.photons * tune_scale))) for s in self.specs]
            with Batch(scaled, absorption=absorption, use_graph=use_graph, device=device) as sample:
                self.tuning = autotune.tune_batch(sample, apply_to=self, device=device)

    def close(self):
        if getattr(self, "_h", None):
            lib().xmc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_launch(self, block: int, blocks_per_cu: int = 0, use_graph: bool = True):
        _check(lib().xmc_batch_set_launch(self._h, block, blocks_per_cu, int(use_graph)), "xmc_batch_set_launch")

    def get_launch(self):
        b, n = C.c_int(), C.c_int()
        _check(lib().xmc_batch_get_launch(self._h, C.byref(b), C.byref(n)), "xmc_batch_get_launch")
        return b.value, n.value

    def run(self, seed: int, stream: Optional[int] = None):
        _check(lib().xmc_batch_run(self._h, seed, stream), "xmc_batch_run")

    def collect(self):
        from .abi import Output
        outs = [OutputHandle(s, absorption=self.absorption) for s in self.specs]
        arr = (Output * len(outs))(*[o.c for o in outs])
        _check(lib().xmc_batch_collect(self._h, arr), "xmc_batch_collect")
        for o, c in zip(outs, arr):
            o.c = c
        return outs

    def stats(self) -> dict:
        s = Stats()
        _check(lib().xmc_batch_stats(self._h, C.byref(s)), "xmc_batch_stats")
        return {k: getattr(s, k) for k, _ in Stats._fields_}
