"""gadget_ng_pm_hip -- Python host mirror of the MI355X PM solver's C ABI.

The product is ``libgng_pm.so`` (HIP kernels for gfx950 behind the C ABI in
``include/gng_pm.h``).  This module binds that ABI with ctypes and mirrors the
reference's solver interface for this path so callers and tests read like the
reference's own:

* :class:`PmSolver` -- ``gadget_ng_pm::PmSolver`` (crates/gadget-ng-pm/src/solver.rs:23-140):
  ``PmSolver(grid_size, box_size)``, public ``plummer_eps``, and
  ``accelerations_for_indices(global_positions, global_masses, eps2, g,
  global_indices, out)`` from the ``GravitySolver`` trait
  (crates/gadget-ng-core/src/gravity.rs:14-24).
* :class:`HipPmSolver` -- ``gadget_ng_hip::HipPmSolver``
  (crates/gadget-ng-hip/src/pm_solver.rs:17-258): ``availability``,
  ``is_available``, ``try_new[_checked]``, ``try_new_with_r_split[_checked]``,
  ``try_accelerations_for_indices`` (raises :class:`HipExecutionError`) and
  ``accelerations_for_indices`` (raises ``RuntimeError`` like the Rust panic).
* :class:`TreePmLongRange` -- the long-range half of ``TreePmSolver``
  (crates/gadget-ng-treepm/src/solver.rs:44-90): r_s = r_split or 2.5 cells.

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_PKG_DIR)  # gadget-ng_amd/
LIB_PATH = os.path.join(_ROOT, "lib", "libgng_pm.so")
CSRC = os.path.join(_ROOT, "csrc")
REPO_ROOT = os.path.dirname(_ROOT)
HEADER = os.path.join(REPO_ROOT, "include", "gng_pm.h")

# the GNG_PM_ABI_VERSION whose struct layouts (GngConfig, GngTuning) these
# bindings mirror
ABI_VERSION = 2

GNG_OK = 0
GNG_ERR_INVALID = 1
GNG_ERR_HIP = 2
GNG_ERR_NO_DEVICE = 3
GNG_ERR_OOM = 4
GNG_ERR_UNSUPPORTED = 5

STAGES = ("bucket", "sort", "deposit", "fft_x_fwd", "fft_y_fwd", "fft_z_green", "fft_y_inv", "fft_x_inv", "interp")

SOURCES = ("gng_pm.hip", "learned.hip", "shortrange.hip", "pm64.hip", "integrate.hip")
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False) -> str:
    """Compile libgng_pm.so for gfx950 in-tree (always recompiles): one hipcc
    per source in parallel, then one link."""
    import tempfile
    os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB_PATH + ".tmp"
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="gng_build_") as work:
        objs, procs = [], []
        for f in SOURCES:
            obj = os.path.join(work, f + ".o")
            objs.append(obj)
            procs.append(subprocess.Popen([hipcc, *compile_flags, "-c", os.path.join(CSRC, f), "-o", obj],
                                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
        errors = []
        for f, pr in zip(SOURCES, procs):
            out, err = pr.communicate()
            if pr.returncode != 0:
                errors.append(f"{f}:\n{out}{err}")
            elif verbose and err:
                print(err)
        if errors:
            raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
        r = subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


class GngConfig(ctypes.Structure):
    _fields_ = [("grid_size", ctypes.c_int32), ("deconvolve", ctypes.c_int32), ("box_size", ctypes.c_double),
                ("r_split", ctypes.c_double), ("plummer_eps", ctypes.c_double), ("device", ctypes.c_int32),
                ("reserved", ctypes.c_int32 * 7)]


class GngTuning(ctypes.Structure):
    _fields_ = [("z_persistent", ctypes.c_int32), ("deposit_wide", ctypes.c_int32),
                ("y_fwd_persistent", ctypes.c_int32), ("y_inv_persistent", ctypes.c_int32),
                ("tuned", ctypes.c_int32), ("interp_sorted", ctypes.c_int32),
                ("plane_chunk_fwd", ctypes.c_int32), ("plane_chunk_inv", ctypes.c_int32),
                ("deposit_split", ctypes.c_int32), ("y_inv_narrow", ctypes.c_int32),
                ("deposit_pair", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {"z_persistent": bool(self.z_persistent), "deposit_threads": 512 if self.deposit_wide else 256,
                "y_fwd_persistent": bool(self.y_fwd_persistent), "y_inv_persistent": bool(self.y_inv_persistent),
                "interp_sorted": bool(self.interp_sorted), "chunk_fwd": int(self.plane_chunk_fwd),
                "chunk_inv": int(self.plane_chunk_inv), "deposit_split": bool(self.deposit_split),
                "y_inv_narrow": bool(self.y_inv_narrow), "deposit_pair": bool(self.deposit_pair),
                "tuned": bool(self.tuned)}


_lib = None


def lib():
    """Load libgng_pm.so (raises if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; "
                               f"g.build()')")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        f32 = ctypes.c_float
        d = ctypes.c_double
        i32 = ctypes.c_int
        L.gng_pm_abi_version.restype = i32
        if L.gng_pm_abi_version() != ABI_VERSION:
            raise RuntimeError(f"{LIB_PATH} has ABI version {L.gng_pm_abi_version()}, this binding expects "
                               f"{ABI_VERSION}: rebuild the library")
        L.gng_last_error_string.restype = ctypes.c_char_p
        L.gng_pm_mesh_supported.argtypes = [i32]
        L.gng_pm_mesh_supported.restype = i32
        L.gng_pm_device_count.restype = i32
        L.gng_pm_create.argtypes = [ctypes.POINTER(GngConfig), ctypes.POINTER(P)]
        L.gng_pm_create.restype = i32
        L.gng_pm_destroy.argtypes = [P]
        L.gng_pm_destroy.restype = None
        L.gng_pm_accelerations_f64.argtypes = [P, P, P, i64, P, i64, d, d, P]
        L.gng_pm_accelerations_f32.argtypes = [P, P, P, P, P, i64, P, i64, d, P, P, P]
        L.gng_pm_solve_device.argtypes = [P, P, P, P, P, i64, P, i64, d, P, P, P, P]
        L.gng_pm_solve_device_f64.argtypes = [P, P, P, i64, P, i64, d, P, P]
        L.gng_pm_solve_device_f64.restype = i32
        L.gng_pm_deposit_order.argtypes = [P, P, P, i64, P, P]
        L.gng_pm_deposit_order.restype = i32
        L.gng_pm_cell_moments.argtypes = [P, P, P]
        L.gng_pm_cell_moments.restype = i32
        L.gng_pm_solve_device_soa64.argtypes = [P, P, P, P, P, i64, P, i64, d, P, P, P, P]
        L.gng_pm_solve_device_soa64.restype = i32
        L.gng_pm_solve_device_noalloc.argtypes = [ctypes.POINTER(GngConfig), P, P, P, P, i64, P, i64, d, P, P, P, P]
        L.gng_pm_solve_device_noalloc.restype = i32
        L.gng_pm_density.argtypes = [P, P]
        L.gng_pm_keep_density.argtypes = [P, i32]
        L.gng_pm_force_meshes.argtypes = [P, P]
        L.gng_pm_set_timing.argtypes = [P, i32]
        L.gng_pm_stage_times.argtypes = [P, P, i32]
        for f in (L.gng_pm_accelerations_f64, L.gng_pm_accelerations_f32, L.gng_pm_solve_device, L.gng_pm_density,
                  L.gng_pm_keep_density, L.gng_pm_force_meshes, L.gng_pm_set_timing, L.gng_pm_stage_times):
            f.restype = i32
        L.gng_dev_alloc.argtypes = [ctypes.POINTER(P), ctypes.c_uint64]
        L.gng_dev_free.argtypes = [P]
        L.gng_dev_copy.argtypes = [P, P, ctypes.c_uint64, i32]
        L.gng_dev_synchronize.argtypes = []
        L.gng_dev_axpy.argtypes = [P, P, i64, f32, P]
        for f in (L.gng_dev_alloc, L.gng_dev_free, L.gng_dev_copy, L.gng_dev_synchronize, L.gng_dev_axpy):
            f.restype = i32
        L.gng_pm_batch_create.argtypes = [i32, ctypes.POINTER(GngConfig), ctypes.POINTER(P)]
        L.gng_pm_batch_create.restype = i32
        L.gng_pm_batch_destroy.argtypes = [P]
        L.gng_pm_batch_destroy.restype = None
        L.gng_pm_batch_size.argtypes = [P]
        L.gng_pm_batch_size.restype = i32
        L.gng_pm_batch_plan.argtypes = [P, i32, ctypes.POINTER(P)]
        L.gng_pm_batch_plan.restype = i32
        L.gng_pm_batch_solve_device.argtypes = [P, P, P, P, P, P, d, P, P, P, P]
        L.gng_pm_batch_solve_device.restype = i32
        L.gng_pm_batch_set_box_tables.argtypes = [P, i32]
        L.gng_pm_batch_set_box_tables.restype = i32
        L.gng_pm_batch_get_box_tables.argtypes = [P]
        L.gng_pm_batch_get_box_tables.restype = i32
        L.gng_pm_batch_tune_launches.argtypes = [P, P, P, P, P, P, d, P, P, P, i32, ctypes.POINTER(i32)]
        L.gng_pm_batch_tune_launches.restype = i32
        L.gng_hilbert_key3.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32]
        L.gng_hilbert_key3.restype = ctypes.c_uint64
        L.gng_pm_autotune.argtypes = [P, i32]
        L.gng_pm_autotune_f64.argtypes = [P, P, P, i64, P, i32]
        L.gng_pm_autotune_f64.restype = i32
        L.gng_pm_get_tuning.argtypes = [P, ctypes.POINTER(GngTuning)]
        L.gng_pm_batch_autotune.argtypes = [P, i32]
        L.gng_pm_set_tuning.argtypes = [P, ctypes.POINTER(GngTuning)]
        for f in (L.gng_pm_autotune, L.gng_pm_get_tuning, L.gng_pm_batch_autotune, L.gng_pm_set_tuning):
            f.restype = i32
        L.gng_pm_keep_potential.argtypes = [P, i32]
        L.gng_pm_potential.argtypes = [P, P]
        L.gng_pm_device_meshes.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P)]
        for f in (L.gng_pm_keep_potential, L.gng_pm_potential, L.gng_pm_device_meshes):
            f.restype = i32
        # short range and learned correction (include/gng_learned.h)
        L.gng_sr_create.argtypes = [i32, ctypes.POINTER(P)]
        L.gng_sr_create.restype = i32
        L.gng_sr_destroy.argtypes = [P]
        L.gng_sr_destroy.restype = None
        L.gng_sr_accelerations.argtypes = [P, P, P, P, P, i64, d, d, d, d, d, i32, P, P, P, P]
        L.gng_sr_accelerations.restype = i32
        L.gng_sr_last_pair_count.argtypes = [P]
        L.gng_sr_last_pair_count.restype = i64
        L.gng_sr_cell_order.argtypes = [P, P, P, P, i64, d, d, P, P]
        L.gng_sr_cell_order.restype = i32
        L.gng_net_param_count.restype = i32
        L.gng_net_create.argtypes = [i32, i32, i32, ctypes.POINTER(P)]
        L.gng_net_destroy.argtypes = [P]
        L.gng_net_destroy.restype = None
        L.gng_net_set_params.argtypes = [P, P]
        L.gng_net_get_params.argtypes = [P, P]
        L.gng_net_buffer_bytes.argtypes = [P, i32]
        L.gng_net_buffer_bytes.restype = i64
        L.gng_net_read.argtypes = [P, i32, P, i64]
        L.gng_net_write.argtypes = [P, i32, P, i64]
        L.gng_net_device_buffer.argtypes = [P, i32, ctypes.POINTER(P)]
        L.gng_net_device_buffer.restype = i32
        L.gng_net_input.argtypes = [P, i32, P, P, f32, f32, P]
        L.gng_net_input_particles.argtypes = [P, i32, P, P, f32, f32, P, P, P, P, i64, d, P]
        L.gng_net_input_moments.argtypes = [P, i32, P, P, f32, f32, P, P]
        L.gng_net_forward.argtypes = [P, i32, P]
        L.gng_net_infer.argtypes = [P, i32, P]
        L.gng_net_infer.restype = i32
        L.gng_net_loss_grad.argtypes = [P, P, P, i64, d, f32, P]
        L.gng_net_backward.argtypes = [P, i32, P]
        L.gng_net_adam.argtypes = [P, f32, f32, f32, f32, P]
        L.gng_net_train_step.argtypes = [P, i32, P, P, i64, d, f32, f32, P]
        L.gng_net_loss.argtypes = [P, ctypes.POINTER(d)]
        L.gng_net_apply.argtypes = [P, i32, P, P, P, i64, d, f32, P, P, P, P]
        L.gng_net_apply_field.argtypes = [P, i32, P, P, P, i64, d, f32, P, d, P, P, P, P]
        for f in (L.gng_net_create, L.gng_net_set_params, L.gng_net_get_params, L.gng_net_read, L.gng_net_write,
                  L.gng_net_input, L.gng_net_input_particles, L.gng_net_input_moments, L.gng_net_forward,
                  L.gng_net_loss_grad,
                  L.gng_net_backward, L.gng_net_adam, L.gng_net_train_step, L.gng_net_loss, L.gng_net_apply,
                  L.gng_net_apply_field):
            f.restype = i32
        L.gng_pm64_create.argtypes = [ctypes.POINTER(GngConfig), ctypes.POINTER(P)]
        L.gng_pm64_destroy.argtypes = [P]
        L.gng_pm64_destroy.restype = None
        L.gng_pm64_accelerations.argtypes = [P, P, P, i64, P, i64, d, P]
        L.gng_pm64_solve_device.argtypes = [P, P, P, i64, P, i64, d, P, P]
        for f in (L.gng_pm64_create, L.gng_pm64_accelerations, L.gng_pm64_solve_device):
            f.restype = i32
        L.gng_pm_power_spectrum.argtypes = [P, P, P, P, P, i64, i32, P, P, P, ctypes.POINTER(i32), P]
        L.gng_pm_power_spectrum_f64.argtypes = [P, P, P, i64, i32, P, P, P, ctypes.POINTER(i32)]
        L.gng_pm_power_spectrum.restype = i32
        L.gng_pm_power_spectrum_f64.restype = i32
        L.gng_pm_get_config.argtypes = [P, ctypes.POINTER(GngConfig)]
        L.gng_pm_get_config.restype = i32
        # device-resident KDK leapfrog (include/gng_integrate.h)
        L.gng_cosmo_advance_a.argtypes = [P, d, d]
        L.gng_cosmo_advance_a.restype = d
        L.gng_cosmo_drift_kick_factors.argtypes = [P, d, d, P]
        L.gng_kdk_create.argtypes = [P, i64, ctypes.POINTER(P)]
        L.gng_kdk_destroy.argtypes = [P]
        L.gng_kdk_destroy.restype = None
        L.gng_kdk_set_short_range.argtypes = [P, P, d, d, d, i32]
        L.gng_kdk_set_state.argtypes = [P, P, P, P]
        L.gng_kdk_set_state_device.argtypes = [P, P, P, P, d, P]
        L.gng_kdk_set_state_device.restype = i32
        L.gng_kdk_get_state.argtypes = [P, P, P, P]
        L.gng_kdk_step.argtypes = [P, d, d, d, d, i32, P]
        L.gng_kdk_run.argtypes = [P, P, P, d, d, i32, i32, P]
        for f in (L.gng_cosmo_drift_kick_factors, L.gng_kdk_create, L.gng_kdk_set_short_range, L.gng_kdk_set_state,
                  L.gng_kdk_get_state, L.gng_kdk_step, L.gng_kdk_run):
            f.restype = i32
        for f in (L.gng_cosmo_hubble, L.gng_cosmo_growth_rate, L.gng_cosmo_growth_factor):
            f.argtypes = [P, d]
            f.restype = d
        # Zel'dovich initial conditions (include/gng_ic.h)
        u64 = ctypes.c_uint64
        L.gng_ic_delta_kspace.argtypes = [P, u64, P, i64, P]
        L.gng_ic_zeldovich.argtypes = [P, u64, P, i64, d, d, i64, i64, P, P]
        L.gng_ic_zeldovich_device.argtypes = [P, u64, P, i64, d, d, i64, i64, P, P, P]
        L.gng_ic_2lpt.argtypes = [P, u64, P, i64, d, d, d, d, i64, i64, P, P]
        L.gng_ic_2lpt_device.argtypes = [P, u64, P, i64, d, d, d, d, i64, i64, P, P, P]
        for f in (L.gng_ic_delta_kspace, L.gng_ic_zeldovich, L.gng_ic_zeldovich_device, L.gng_ic_2lpt,
                  L.gng_ic_2lpt_device):
            f.restype = i32
        L.hip_pm_create.argtypes = [i32, f32]
        L.hip_pm_create.restype = P
        L.hip_pm_destroy.argtypes = [P]
        L.hip_pm_destroy.restype = None
        L.hip_pm_solve.argtypes = [P, P, P, P, P, P, P, P, i32, f32, f32, f32]
        L.hip_pm_solve.restype = i32
        _lib = L
    return _lib


def hilbert_key3(x: int, y: int, z: int, bits: int) -> int:
    """3-D Hilbert key used to order the deposit tiles (host function, no GPU)."""
    return int(lib().gng_hilbert_key3(x, y, z, bits))


def last_error() -> str:
    return lib().gng_last_error_string().decode(errors="replace")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class HipExecutionError(RuntimeError):
    """A solve returned a nonzero code (crates/gadget-ng-hip/src/availability.rs)."""

    def __init__(self, kernel: str, code: int, message: str = ""):
        super().__init__(f"{kernel} failed with code {code}: {message}")
        self.kernel = kernel
        self.code = code


@dataclass
class HipAvailability:
    """Build / runtime availability, as HipAvailability (availability.rs:7-40)."""
    compiled: bool
    runtime_available: bool
    reason: str

    def is_available(self) -> bool:
        return self.compiled and self.runtime_available


class HipUnavailable(RuntimeError):
    def __init__(self, availability: HipAvailability):
        super().__init__(f"HIP PM solver unavailable: {availability.reason}")
        self.availability = availability


class Plan:
    """Owns one gng_pm_plan (device meshes for one mesh size and filter set)."""

    def __init__(self, grid_size: int, box_size: float, r_split: float = 0.0, plummer_eps: float = 0.0,
                 deconvolve: bool = False, device: int = -1, keep_density: bool = False,
                 keep_potential: bool = False):
        cfg = GngConfig()
        cfg.grid_size = int(grid_size)
        cfg.deconvolve = int(bool(deconvolve))
        cfg.box_size = float(box_size)
        cfg.r_split = float(r_split)
        cfg.plummer_eps = float(plummer_eps)
        cfg.device = int(device)
        h = ctypes.c_void_p()
        rc = lib().gng_pm_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm_create", rc, last_error())
        self._h = h
        if keep_density:
            self._check(lib().gng_pm_keep_density(h, 1), "gng_pm_keep_density")
        if keep_potential:
            self._check(lib().gng_pm_keep_potential(h, 1), "gng_pm_keep_potential")
        self.grid_size = int(grid_size)
        self.box_size = float(box_size)
        self.r_split = float(r_split)
        self.plummer_eps = float(plummer_eps)
        self.deconvolve = bool(deconvolve)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().gng_pm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc != GNG_OK:
            raise HipExecutionError(what, rc, last_error())

    def accelerations_f64(self, pos: np.ndarray, mass: np.ndarray, g: float, idx=None,
                          eps2: float = 0.0, out: Optional[np.ndarray] = None) -> np.ndarray:
        """Accelerations (k x 3 float64) of pos[idx] (all n without idx).  out,
        when given, is a C-contiguous float64 (k, 3) array the result is
        written into and returned, as the reference's solvers fill the
        caller's `out: &mut [Vec3]` (a caller that keeps it across steps
        saves the page faults of a fresh array on every call)."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        n = pos.shape[0]
        if mass.shape[0] != n:
            raise ValueError("positions and masses differ in length")
        if idx is None:
            k, ip = n, None
        else:
            idx = np.ascontiguousarray(idx, dtype=np.int64)
            k, ip = idx.shape[0], _ptr(idx)
        if out is None:
            out = np.empty((k, 3), dtype=np.float64)
        elif (not isinstance(out, np.ndarray) or out.dtype != np.float64 or out.shape != (k, 3)
              or not out.flags.c_contiguous or not out.flags.writeable):
            raise ValueError(f"out must be a writeable C-contiguous float64 array of shape ({k}, 3)")
        self._check(lib().gng_pm_accelerations_f64(self._h, _ptr(pos), _ptr(mass), n, ip, k, float(eps2), float(g),
                                                   _ptr(out)), "gng_pm_accelerations_f64")
        return out

    def accelerations_f32(self, x, y, z, m, g: float, idx=None, out: Optional[np.ndarray] = None):
        """Accelerations (3 x k float32, one row per axis) of particles idx
        (all n without idx); out, when given, is a writeable C-contiguous
        float32 (3, k) array the result is written into and returned."""
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (x, y, z, m)]
        n = arrs[0].shape[0]
        if idx is None:
            k, ip = n, None
        else:
            idx = np.ascontiguousarray(idx, dtype=np.int64)
            k, ip = idx.shape[0], _ptr(idx)
        if out is None:
            out = np.empty((3, k), dtype=np.float32)
        elif (not isinstance(out, np.ndarray) or out.dtype != np.float32 or out.shape != (3, k)
              or not out.flags.c_contiguous or not out.flags.writeable):
            raise ValueError(f"out must be a writeable C-contiguous float32 array of shape (3, {k})")
        self._check(lib().gng_pm_accelerations_f32(self._h, *(_ptr(a) for a in arrs), n, ip, k, float(g),
                                                   _ptr(out[0]), _ptr(out[1]), _ptr(out[2])),
                    "gng_pm_accelerations_f32")
        return out

    def solve_device(self, x_ptr: int, y_ptr: int, z_ptr: int, m_ptr: int, n: int, ax_ptr: int, ay_ptr: int,
                     az_ptr: int, g: float, stream: int = 0, idx_ptr: Optional[int] = None,
                     k: Optional[int] = None):
        """Enqueue a solve on raw device pointers (e.g. torch tensors' data_ptr())."""
        k = n if k is None else k
        self._check(lib().gng_pm_solve_device(self._h, x_ptr, y_ptr, z_ptr, m_ptr, n, idx_ptr, k, float(g), ax_ptr,
                                              ay_ptr, az_ptr, stream or None), "gng_pm_solve_device")

    def solve_device_f64(self, pos_ptr: int, m_ptr: int, n: int, out_ptr: int, g: float, stream: int = 0,
                         idx_ptr: Optional[int] = None, k: Optional[int] = None):
        """Enqueue a solve on device Vec3 arrays: n x 3 f64 positions, n f64
        masses, k x 3 f64 accelerations out."""
        k = n if k is None else k
        self._check(lib().gng_pm_solve_device_f64(self._h, pos_ptr, m_ptr, n, idx_ptr, k, float(g), out_ptr,
                                                  stream or None), "gng_pm_solve_device_f64")

    def deposit_order(self, pos, mass) -> np.ndarray:
        """The deposit's sorted particle order for (n, 3) f64 positions and n
        masses (gng_pm_deposit_order): uint32 index of the particle in each
        sorted slot."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        dpos = DeviceArray.from_host(pos)
        dm = DeviceArray.from_host(np.ascontiguousarray(mass, dtype=np.float64))
        out = DeviceArray((n,), np.uint32)
        self._check(lib().gng_pm_deposit_order(self._h, dpos.ptr, dm.ptr, n, out.ptr, None), "gng_pm_deposit_order")
        device_synchronize()
        return out.numpy()

    def solve_device_soa64(self, x_ptr: int, y_ptr: int, z_ptr: int, m_ptr: int, n: int, ax_ptr: int, ay_ptr: int,
                           az_ptr: int, g: float, stream: int = 0, idx_ptr: Optional[int] = None,
                           k: Optional[int] = None):
        """Enqueue a solve on device SoA f64 positions, f32 masses, f32 accelerations."""
        k = n if k is None else k
        self._check(lib().gng_pm_solve_device_soa64(self._h, x_ptr, y_ptr, z_ptr, m_ptr, n, idx_ptr, k, float(g),
                                                    ax_ptr, ay_ptr, az_ptr, stream or None),
                    "gng_pm_solve_device_soa64")

    def density(self) -> np.ndarray:
        """Mass per cell of the last solve (needs keep_density=True)."""
        nm = self.grid_size
        out = np.empty((nm, nm, nm), dtype=np.float64)
        self._check(lib().gng_pm_density(self._h, _ptr(out)), "gng_pm_density")
        return out

    def autotune(self, reps: int = 5) -> dict:
        """Time the compiled launch variants on synthetic particles and keep the fastest."""
        self._check(lib().gng_pm_autotune(self._h, int(reps)), "gng_pm_autotune")
        return self.tuning()

    def autotune_f64(self, pos_ptr: int, m_ptr: int, n: int, out_ptr: int, reps: int = 5) -> dict:
        """The autotuner timed on the caller's device Vec3 particles."""
        self._check(lib().gng_pm_autotune_f64(self._h, pos_ptr, m_ptr, int(n), out_ptr, int(reps)),
                    "gng_pm_autotune_f64")
        return self.tuning()

    def set_tuning(self, z_persistent: bool = False, deposit_threads: int = 512, y_fwd_persistent: bool = False,
                   y_inv_persistent: bool = False, interp_sorted: bool = False, chunk_fwd: int = 0,
                   chunk_inv: int = 0, deposit_split: bool = True, y_inv_narrow: bool = True,
                   deposit_pair: bool = False) -> None:
        if deposit_threads not in (256, 512):
            raise ValueError("deposit_threads must be 256 or 512")
        t = GngTuning()
        t.z_persistent = int(bool(z_persistent))
        t.deposit_wide = int(deposit_threads == 512)
        t.y_fwd_persistent = int(bool(y_fwd_persistent))
        t.y_inv_persistent = int(bool(y_inv_persistent))
        t.interp_sorted = int(bool(interp_sorted))
        t.plane_chunk_fwd = int(chunk_fwd)
        t.plane_chunk_inv = int(chunk_inv)
        t.deposit_split = int(bool(deposit_split))
        t.y_inv_narrow = int(bool(y_inv_narrow))
        t.deposit_pair = int(bool(deposit_pair))
        self._check(lib().gng_pm_set_tuning(self._h, ctypes.byref(t)), "gng_pm_set_tuning")

    def tuning(self) -> dict:
        t = GngTuning()
        self._check(lib().gng_pm_get_tuning(self._h, ctypes.byref(t)), "gng_pm_get_tuning")
        return t.as_dict()

    def potential(self) -> np.ndarray:
        """Mesh potential of the last solve (needs keep_potential=True)."""
        nm = self.grid_size
        out = np.empty((nm, nm, nm), dtype=np.float64)
        self._check(lib().gng_pm_potential(self._h, _ptr(out)), "gng_pm_potential")
        return out

    def cell_moments_device(self, mom_ptr: int, stream: int = 0) -> None:
        """Enqueue the per-cell particle moments of the last solve into a device
        buffer of 5 doubles per cell (gng_pm_cell_moments)."""
        self._check(lib().gng_pm_cell_moments(self._h, mom_ptr, stream or None), "gng_pm_cell_moments")

    def cell_moments(self) -> np.ndarray:
        """(nm, nm, nm, 5) f64 moments of the last solve's particles."""
        nm = self.grid_size
        out = DeviceArray((nm, nm, nm, 5), np.float64)
        self.cell_moments_device(out.ptr)
        device_synchronize()
        return out.numpy()

    def device_meshes(self):
        """(mass-per-cell, potential) device pointers of the last solve."""
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        self._check(lib().gng_pm_device_meshes(self._h, ctypes.byref(a), ctypes.byref(b)), "gng_pm_device_meshes")
        return int(a.value), int(b.value)

    def force_meshes(self) -> np.ndarray:
        """(3, nm, nm, nm) f32 force meshes of the last solve."""
        nm = self.grid_size
        out = np.empty((3, nm, nm, nm), dtype=np.float32)
        self._check(lib().gng_pm_force_meshes(self._h, _ptr(out)), "gng_pm_force_meshes")
        return out

    def set_timing(self, on: bool = True):
        self._check(lib().gng_pm_set_timing(self._h, int(on)), "gng_pm_set_timing")

    def stage_times(self) -> dict:
        """Mean per-stage milliseconds over the solves recorded since set_timing(True),
        plus 'total' (first to last event of a solve)."""
        ms = np.zeros(16, dtype=np.float32)
        n = lib().gng_pm_stage_times(self._h, _ptr(ms), 16)
        if n < 0:
            raise HipExecutionError("gng_pm_stage_times", -n, last_error())
        out = {STAGES[i]: float(ms[i]) for i in range(min(n, len(STAGES)))}
        out["total"] = float(ms[n])
        return out


class Plan64:
    """The FP64 GPU form of the reference algorithm (the baseline the FP32
    solver is compared with): same configuration as Plan, powers of two only."""

    def __init__(self, grid_size: int, box_size: float, r_split: float = 0.0, plummer_eps: float = 0.0,
                 deconvolve: bool = False, device: int = -1):
        cfg = GngConfig()
        cfg.grid_size = int(grid_size)
        cfg.deconvolve = int(bool(deconvolve))
        cfg.box_size = float(box_size)
        cfg.r_split = float(r_split)
        cfg.plummer_eps = float(plummer_eps)
        cfg.device = int(device)
        h = ctypes.c_void_p()
        rc = lib().gng_pm64_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm64_create", rc, last_error())
        self._h = h
        self.grid_size = int(grid_size)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().gng_pm64_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def accelerations(self, pos, mass, g: float, idx=None) -> np.ndarray:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        n = pos.shape[0]
        if idx is None:
            k, ip = n, None
        else:
            idx = np.ascontiguousarray(idx, dtype=np.int64)
            k, ip = idx.shape[0], _ptr(idx)
        out = np.empty((k, 3), dtype=np.float64)
        rc = lib().gng_pm64_accelerations(self._h, _ptr(pos), _ptr(mass), n, ip, k, float(g), _ptr(out))
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm64_accelerations", rc, last_error())
        return out

    def solve_device(self, pos_ptr: int, mass_ptr: int, n: int, out_ptr: int, g: float, stream: int = 0) -> None:
        rc = lib().gng_pm64_solve_device(self._h, pos_ptr, mass_ptr, int(n), None, int(n), float(g), out_ptr,
                                         stream or None)
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm64_solve_device", rc, last_error())


NOALLOC_MAX_GRID = 64
NOALLOC_MAX_PARTICLES = 1 << 19


def solve_device_noalloc(grid_size: int, box_size: float, x_ptr: int, y_ptr: int, z_ptr: int, m_ptr: int, n: int,
                         ax_ptr: int, ay_ptr: int, az_ptr: int, g: float, r_split: float = 0.0,
                         plummer_eps: float = 0.0, deconvolve: bool = False, device: int = -1, stream: int = 0,
                         idx_ptr: Optional[int] = None, k: Optional[int] = None) -> None:
    """gng_pm_solve_device_noalloc: a plan-free solve on device pointers whose
    work buffers live in the library's static per-GPU scratch (grid_size <=
    64, n <= 2^19); calls on one GPU must not overlap."""
    cfg = GngConfig()
    cfg.grid_size = int(grid_size)
    cfg.deconvolve = int(bool(deconvolve))
    cfg.box_size = float(box_size)
    cfg.r_split = float(r_split)
    cfg.plummer_eps = float(plummer_eps)
    cfg.device = int(device)
    k = n if k is None else k
    rc = lib().gng_pm_solve_device_noalloc(ctypes.byref(cfg), x_ptr, y_ptr, z_ptr, m_ptr, int(n), idx_ptr, int(k),
                                           float(g), ax_ptr, ay_ptr, az_ptr, stream or None)
    if rc != GNG_OK:
        raise HipExecutionError("gng_pm_solve_device_noalloc", rc, last_error())


class DeviceArray:
    """A device buffer owned through the library's own HIP runtime (no torch needed)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        rc = lib().gng_dev_alloc(ctypes.byref(p), max(self.nbytes, 1))
        if rc != GNG_OK:
            raise HipExecutionError("gng_dev_alloc", rc, last_error())
        self._p = p

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self) -> int:
        return int(self._p.value or 0)

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        rc = lib().gng_dev_copy(self._p, _ptr(a), self.nbytes, 0)
        if rc != GNG_OK:
            raise HipExecutionError("gng_dev_copy", rc, last_error())

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        rc = lib().gng_dev_copy(_ptr(out), self._p, self.nbytes, 1)
        if rc != GNG_OK:
            raise HipExecutionError("gng_dev_copy", rc, last_error())
        return out

    def free(self) -> None:
        if getattr(self, "_p", None) is not None and self._p.value:
            lib().gng_dev_free(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_copy(dst: DeviceArray, src: DeviceArray) -> None:
    """Device-to-device copy of src into dst (same byte size)."""
    assert dst.nbytes == src.nbytes
    rc = lib().gng_dev_copy(dst._p, src._p, src.nbytes, 2)
    if rc != GNG_OK:
        raise HipExecutionError("gng_dev_copy", rc, last_error())


def device_synchronize() -> None:
    rc = lib().gng_dev_synchronize()
    if rc != GNG_OK:
        raise HipExecutionError("gng_dev_synchronize", rc, last_error())


class Batch:
    """An ensemble of independent periodic boxes solved together (one HIP graph)."""

    def __init__(self, boxes, device: int = -1):
        """boxes: sequence of dicts with grid_size, box_size and optional r_split,
        plummer_eps, deconvolve."""
        cfgs = (GngConfig * len(boxes))()
        for c, b in zip(cfgs, boxes):
            c.grid_size = int(b["grid_size"])
            c.box_size = float(b["box_size"])
            c.r_split = float(b.get("r_split", 0.0))
            c.plummer_eps = float(b.get("plummer_eps", 0.0))
            c.deconvolve = int(bool(b.get("deconvolve", False)))
            c.device = int(device)
        h = ctypes.c_void_p()
        rc = lib().gng_pm_batch_create(len(boxes), cfgs, ctypes.byref(h))
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm_batch_create", rc, last_error())
        self._h = h
        self.n_boxes = len(boxes)

    def autotune(self, reps: int = 3) -> None:
        rc = lib().gng_pm_batch_autotune(self._h, int(reps))
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm_batch_autotune", rc, last_error())

    def set_box_tables(self, boxes_per_launch) -> None:
        """k >= 1: every stage one launch over at most k boxes of one mesh
        side (True: one launch per side); 0 or False: one pipeline per box."""
        k = (1 << 30) if boxes_per_launch is True else int(boxes_per_launch)
        rc = lib().gng_pm_batch_set_box_tables(self._h, k)
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm_batch_set_box_tables", rc, last_error())

    def box_tables(self) -> int:
        return int(lib().gng_pm_batch_get_box_tables(self._h))

    def tune_launches(self, x: int, y: int, z: int, m: int, offsets, g: float, ax: int, ay: int, az: int,
                      reps: int = 3) -> int:
        """Time the launch splits on these particles and keep the fastest (returns k)."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        k = ctypes.c_int32()
        rc = lib().gng_pm_batch_tune_launches(self._h, x, y, z, m, _ptr(off), float(g), ax, ay, az, int(reps),
                                              ctypes.byref(k))
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm_batch_tune_launches", rc, last_error())
        return int(k.value)

    def tuning(self, i: int) -> dict:
        p = ctypes.c_void_p()
        rc = lib().gng_pm_batch_plan(self._h, int(i), ctypes.byref(p))
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm_batch_plan", rc, last_error())
        t = GngTuning()
        rc = lib().gng_pm_get_tuning(p, ctypes.byref(t))
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm_get_tuning", rc, last_error())
        return t.as_dict()

    def solve_device(self, x: int, y: int, z: int, m: int, offsets, g: float, ax: int, ay: int, az: int,
                     stream: int = 0):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        assert off.shape[0] == self.n_boxes + 1
        rc = lib().gng_pm_batch_solve_device(self._h, x, y, z, m, _ptr(off), float(g), ax, ay, az, stream or None)
        if rc != GNG_OK:
            raise HipExecutionError("gng_pm_batch_solve_device", rc, last_error())

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().gng_pm_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _as_positions(global_positions) -> np.ndarray:
    return np.ascontiguousarray(global_positions, dtype=np.float64).reshape(-1, 3)


def _fills_directly(out, k: int) -> bool:
    """out can receive the accelerations in place (no temporary and copy)."""
    return (isinstance(out, np.ndarray) and out.dtype == np.float64 and out.shape == (k, 3)
            and out.flags.c_contiguous and out.flags.writeable)


@dataclass
class PmSolver:
    """Drop-in for gadget_ng_pm::PmSolver on one GPU (solver.rs:23-58).

    ``modified_gravity`` (f(R)) is not part of this solver; setting it raises.
    ``deconvolve`` is an extension (False = reference semantics).
    """
    grid_size: int
    box_size: float
    plummer_eps: Optional[float] = None
    deconvolve: bool = False
    modified_gravity: Optional[object] = None
    _plans: dict = field(default_factory=dict, repr=False)

    def _plan(self, r_split: float = 0.0) -> Plan:
        if self.modified_gravity is not None:
            raise NotImplementedError("f(R) modified gravity is not implemented by the GPU PM solver")
        eps = self.plummer_eps if (self.plummer_eps is not None and self.plummer_eps > 0.0) else 0.0
        key = (self.grid_size, self.box_size, float(r_split), eps, self.deconvolve)
        p = self._plans.get(key)
        if p is None:
            p = Plan(self.grid_size, self.box_size, r_split=r_split, plummer_eps=eps, deconvolve=self.deconvolve)
            self._plans[key] = p
        return p

    def accelerations_for_indices(self, global_positions, global_masses, eps2: float, g: float, global_indices,
                                  out: Optional[np.ndarray] = None) -> np.ndarray:
        """GravitySolver::accelerations_for_indices: density from all N, output for the K indices."""
        idx = np.asarray(global_indices, dtype=np.int64)
        if out is not None and out.shape[0] != idx.shape[0]:
            raise AssertionError("out and global_indices differ in length")  # solver.rs:83
        pos = _as_positions(global_positions)
        if _fills_directly(out, idx.shape[0]):
            return self._plan().accelerations_f64(pos, global_masses, g, idx, eps2, out=out)
        res = self._plan().accelerations_f64(pos, global_masses, g, idx, eps2)
        if out is not None:
            out[...] = res
            return out
        return res


@dataclass
class TreePmLongRange:
    """Long-range half of gadget_ng_treepm::TreePmSolver (treepm/src/solver.rs:44-90)."""
    grid_size: int
    box_size: float
    r_split: float = 0.0
    _pm: Optional[PmSolver] = field(default=None, repr=False)

    def effective_r_split(self) -> float:
        """treepm/src/solver.rs:56-62: r_split > 0 or 2.5 cells."""
        return self.r_split if self.r_split > 0.0 else 2.5 * self.box_size / self.grid_size

    def long_range_accelerations_for_indices(self, global_positions, global_masses, eps2: float, g: float,
                                             global_indices) -> np.ndarray:
        if self._pm is None:
            self._pm = PmSolver(self.grid_size, self.box_size)
        idx = np.asarray(global_indices, dtype=np.int64)
        return self._pm._plan(self.effective_r_split()).accelerations_f64(
            _as_positions(global_positions), global_masses, g, idx, eps2)


class HipPmSolver:
    """Mirror of gadget_ng_hip::HipPmSolver (crates/gadget-ng-hip/src/pm_solver.rs:17-258)."""

    def __init__(self, grid_size: int, box_size: float, r_split: float, plan: Plan):
        self._grid_size = int(grid_size)
        self._r_split = float(r_split)
        self._plan = plan

    @staticmethod
    def availability() -> HipAvailability:
        if not os.path.exists(LIB_PATH):
            return HipAvailability(False, False, "libgng_pm.so not built")
        try:
            L = lib()
        except OSError as e:
            return HipAvailability(False, False, f"cannot load libgng_pm.so: {e}")
        if L.gng_pm_device_count() <= 0:
            return HipAvailability(True, False, "no HIP device")
        h = L.hip_pm_create(8, ctypes.c_float(1.0))
        if not h:
            return HipAvailability(True, False, "could not create a runtime handle: " + last_error())
        L.hip_pm_destroy(h)
        return HipAvailability(True, True, "ok")

    @classmethod
    def is_available(cls) -> bool:
        return cls.availability().is_available()

    @classmethod
    def try_new(cls, grid_size: int, box_size: float):
        try:
            return cls.try_new_checked(grid_size, box_size)
        except HipUnavailable:
            return None

    @classmethod
    def try_new_checked(cls, grid_size: int, box_size: float):
        return cls.try_new_with_r_split_checked(grid_size, box_size, 0.0)

    @classmethod
    def try_new_with_r_split(cls, grid_size: int, box_size: float, r_split: float):
        try:
            return cls.try_new_with_r_split_checked(grid_size, box_size, r_split)
        except HipUnavailable:
            return None

    @classmethod
    def try_new_with_r_split_checked(cls, grid_size: int, box_size: float, r_split: float):
        av = cls.availability()
        if not av.is_available():
            raise HipUnavailable(av)
        try:
            plan = Plan(grid_size, box_size, r_split=max(float(r_split), 0.0))
        except HipExecutionError as e:
            raise HipUnavailable(HipAvailability(True, False, str(e))) from e
        return cls(grid_size, box_size, r_split, plan)

    def r_split(self) -> float:
        return self._r_split

    def grid_size(self) -> int:
        return self._grid_size

    def try_accelerations_for_indices(self, global_positions, global_masses, eps2: float, g: float,
                                      global_indices, out: np.ndarray) -> None:
        idx = np.asarray(global_indices, dtype=np.int64)
        assert idx.shape[0] == out.shape[0]
        if idx.shape[0] == 0:
            return
        pos = _as_positions(global_positions)
        if _fills_directly(out, idx.shape[0]):
            self._plan.accelerations_f64(pos, global_masses, g, idx, eps2, out=out)
        else:
            out[...] = self._plan.accelerations_f64(pos, global_masses, g, idx, eps2)

    def accelerations_for_indices(self, global_positions, global_masses, eps2: float, g: float, global_indices,
                                  out: np.ndarray) -> None:
        try:
            self.try_accelerations_for_indices(global_positions, global_masses, eps2, g, global_indices, out)
        except HipExecutionError as e:  # the Rust impl panics (pm_solver.rs:255)
            raise RuntimeError(f"HipPmSolver::accelerations_for_indices failed: {e}") from e
