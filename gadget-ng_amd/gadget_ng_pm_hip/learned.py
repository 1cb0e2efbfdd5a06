"""TreePM on the GPU (long range + short range) and the learned sub-cell
correction, on top of libgng_pm.so (include/gng_learned.h).

* :class:`ShortRange`     direct-sum short-range force (exact or reference split);
* :class:`TreePmSolver`   mirror of crates/gadget-ng-treepm/src/solver.rs:
                          TreePmSolver (PM long range with the Gaussian filter
                          + short range, r_cut = 5 r_s, solver.rs:56-107);
* :class:`Net`            the 3-layer BF16 convolution network (weights,
                          buffers, forward / loss / backward / Adam);
* :class:`LearnedCorrection`  data generation (PM force, TreePM target, network
                          input), training over seeded boxes and evaluation
                          against TreePM on held-out seeds.

Everything runs through the HIP library; nothing here falls back to the CPU.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import (GNG_OK, DeviceArray, HipExecutionError, Plan, _ptr, device_synchronize, last_error, lib)
from . import workloads

SR_EXACT = 0
SR_REFERENCE = 1

BUFFERS = {"x0": 0, "z1": 1, "x1": 2, "z2": 3, "x2": 4, "y": 5, "gy": 6, "dz3": 7, "dz2": 8, "dz1": 9,
           "grad": 10, "params": 11}
# (dtype, channels) of the mesh buffers
_MESH = {"x0": ("bf16", 8), "z1": ("f32", 16), "x1": ("bf16", 16), "z2": ("f32", 16), "x2": ("bf16", 16),
         "y": ("f32", 4), "gy": ("f32", 4), "dz3": ("bf16", 4), "dz2": ("bf16", 16), "dz1": ("bf16", 16)}
LAYERS = ((8, 16), (16, 16), (16, 3))
X0_CHANNELS = LAYERS[0][0]


def _check(rc: int, what: str) -> None:
    if rc != GNG_OK:
        raise HipExecutionError(what, rc, last_error())


def bf16_to_f32(bits: np.ndarray) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)


def f32_to_bf16(a: np.ndarray) -> np.ndarray:
    """Round to nearest even bfloat16 bit patterns."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def pack_params(p: Dict[str, np.ndarray]) -> np.ndarray:
    """Flat float32 parameter vector in the library's order w1 b1 w2 b2 w3 b3."""
    parts = []
    for k, (ci, co) in enumerate(LAYERS, start=1):
        parts.append(np.asarray(p[f"w{k}"], dtype=np.float32).reshape(27 * ci * co))
        parts.append(np.asarray(p[f"b{k}"], dtype=np.float32).reshape(co))
    return np.concatenate(parts)


def unpack_params(flat: np.ndarray) -> Dict[str, np.ndarray]:
    out, o = {}, 0
    for k, (ci, co) in enumerate(LAYERS, start=1):
        out[f"w{k}"] = np.array(flat[o:o + 27 * ci * co], dtype=np.float32).reshape(27, ci, co)
        o += 27 * ci * co
        out[f"b{k}"] = np.array(flat[o:o + co], dtype=np.float32)
        o += co
    return out


def init_params(seed: int = 1234, scale: float = 1.0) -> Dict[str, np.ndarray]:
    """He-normal weights (std scale*sqrt(2/fan_in)), the last layer at a tenth."""
    g = np.random.default_rng(seed)
    p = {}
    for k, (ci, co) in enumerate(LAYERS, start=1):
        std = scale * math.sqrt(2.0 / (27 * ci)) * (0.1 if k == 3 else 1.0)
        p[f"w{k}"] = g.normal(0.0, std, (27, ci, co)).astype(np.float32)
        p[f"b{k}"] = np.zeros(co, dtype=np.float32)
    return p


# ---------------------------------------------------------------------------
# short range and TreePM
# ---------------------------------------------------------------------------

class ShortRange:
    """Direct-sum short-range accelerations on the GPU (short_range.rs:232-398)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().gng_sr_create(int(device), ctypes.byref(h)), "gng_sr_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().gng_sr_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def accelerations_device(self, x: int, y: int, z: int, m: int, n: int, box_size: float, r_split: float,
                             r_cut: float, eps2: float, g: float, kernel: int, ax: int, ay: int, az: int,
                             stream: int = 0) -> None:
        _check(lib().gng_sr_accelerations(self._h, x, y, z, m, int(n), float(box_size), float(r_split),
                                          float(r_cut), float(eps2), float(g), int(kernel), ax, ay, az,
                                          stream or None), "gng_sr_accelerations")

    def accelerations(self, pos, mass, box_size: float, r_split: float, r_cut: Optional[float] = None,
                      eps2: float = 0.0, g: float = 1.0, kernel: str = "exact") -> np.ndarray:
        """Host convenience: float32 copies in, (N, 3) float64 out."""
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        r_cut = 5.0 * r_split if r_cut is None else r_cut
        d = [DeviceArray.from_host(np.ascontiguousarray(pos[:, a], dtype=np.float32)) for a in range(3)]
        dm = DeviceArray.from_host(np.ascontiguousarray(mass, dtype=np.float32))
        out = [DeviceArray((n,), np.float32) for _ in range(3)]
        self.accelerations_device(d[0].ptr, d[1].ptr, d[2].ptr, dm.ptr, n, box_size, r_split, r_cut, eps2, g,
                                  SR_EXACT if kernel == "exact" else SR_REFERENCE, out[0].ptr, out[1].ptr,
                                  out[2].ptr)
        device_synchronize()
        return np.stack([o.numpy() for o in out], axis=1).astype(np.float64)

    def last_pair_count(self) -> int:
        return int(lib().gng_sr_last_pair_count(self._h))

    def cell_order(self, pos, box_size: float, r_cut: float) -> np.ndarray:
        """The stable cell-sorted permutation the pair sums visit (uint32)."""
        pos = np.asarray(pos).reshape(-1, 3)
        n = pos.shape[0]
        d = [DeviceArray.from_host(np.ascontiguousarray(pos[:, a], dtype=np.float32)) for a in range(3)]
        out = DeviceArray((n,), np.uint32)
        _check(lib().gng_sr_cell_order(self._h, d[0].ptr, d[1].ptr, d[2].ptr, int(n), float(box_size),
                                       float(r_cut), out.ptr, None), "gng_sr_cell_order")
        device_synchronize()
        return out.numpy()


class TreePmSolver:
    """TreePmSolver (crates/gadget-ng-treepm/src/solver.rs) on one GPU.

    accelerations = PM long range (filter exp(-k^2 r_s^2 / 2), r_s = r_split or
    2.5 cells, solver.rs:56-62) + short range summed directly within r_cut =
    5 r_s (solver.rs:80).  kernel="exact" uses the split that complements the
    filter exactly (long + short = Newton); kernel="reference" reproduces the
    reference's erfc-only factor (short_range.rs:103-110).  deconvolve divides
    the long-range Green's function by the CIC window squared (an extension).
    """

    def __init__(self, grid_size: int, box_size: float, r_split: float = 0.0, kernel: str = "exact",
                 deconvolve: bool = False, device: int = 0):
        self.grid_size = int(grid_size)
        self.box_size = float(box_size)
        self.r_split = float(r_split) if r_split > 0 else 2.5 * self.box_size / self.grid_size
        self.r_cut = 5.0 * self.r_split
        self.kernel = SR_EXACT if kernel == "exact" else SR_REFERENCE
        self.long_range = Plan(self.grid_size, self.box_size, r_split=self.r_split, deconvolve=deconvolve,
                               device=device)
        self.short_range = ShortRange(device=max(device, 0))

    def accelerations_device(self, x: int, y: int, z: int, m: int, n: int, g: float, eps2: float,
                             ax: int, ay: int, az: int, scratch) -> None:
        """Device-resident total force; scratch = three float32 DeviceArrays of n."""
        self.long_range.solve_device(x, y, z, m, n, ax, ay, az, g)
        self.short_range.accelerations_device(x, y, z, m, n, self.box_size, self.r_split, self.r_cut, eps2, g,
                                              self.kernel, scratch[0].ptr, scratch[1].ptr, scratch[2].ptr)
        _axpy_device([ax, ay, az], scratch, n)

    def accelerations_for_indices(self, global_positions, global_masses, eps2: float, g: float,
                                  global_indices) -> np.ndarray:
        pos = np.asarray(global_positions, dtype=np.float64).reshape(-1, 3)
        idx = np.asarray(global_indices, dtype=np.int64)
        n = pos.shape[0]
        d = [DeviceArray.from_host(np.ascontiguousarray(pos[:, a], dtype=np.float32)) for a in range(3)]
        dm = DeviceArray.from_host(np.ascontiguousarray(global_masses, dtype=np.float32))
        out = [DeviceArray((n,), np.float32) for _ in range(3)]
        tmp = [DeviceArray((n,), np.float32) for _ in range(3)]
        self.accelerations_device(d[0].ptr, d[1].ptr, d[2].ptr, dm.ptr, n, g, eps2, out[0].ptr, out[1].ptr,
                                  out[2].ptr, tmp)
        device_synchronize()
        a = np.stack([o.numpy() for o in out], axis=1).astype(np.float64)
        return a[idx]


def _axpy_device(dst_ptrs, src_arrays, n: int, a: float = 1.0) -> None:
    """dst += a * src for three float32 device arrays."""
    for dp, s in zip(dst_ptrs, src_arrays):
        _check(lib().gng_dev_axpy(dp, s.ptr if isinstance(s, DeviceArray) else s, int(n), float(a), None),
               "gng_dev_axpy")


# ---------------------------------------------------------------------------
# the network
# ---------------------------------------------------------------------------

class Net:
    """One gng_net: weights plus buffers for max_boxes meshes of grid_size^3."""

    def __init__(self, grid_size: int, max_boxes: int = 1, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().gng_net_create(int(grid_size), int(max_boxes), int(device), ctypes.byref(h)), "gng_net_create")
        self._h = h
        self.grid_size = int(grid_size)
        self.max_boxes = int(max_boxes)
        self.M = self.grid_size ** 3

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().gng_net_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def param_count() -> int:
        return int(lib().gng_net_param_count())

    def set_params(self, p) -> None:
        flat = pack_params(p) if isinstance(p, dict) else np.ascontiguousarray(p, dtype=np.float32)
        assert flat.shape[0] == self.param_count()
        _check(lib().gng_net_set_params(self._h, _ptr(flat)), "gng_net_set_params")

    def params(self) -> Dict[str, np.ndarray]:
        flat = np.empty(self.param_count(), dtype=np.float32)
        _check(lib().gng_net_get_params(self._h, _ptr(flat)), "gng_net_get_params")
        return unpack_params(flat)

    def read(self, name: str, boxes: Optional[int] = None) -> np.ndarray:
        """A buffer as float32 (bf16 widened): meshes as [box, nz, ny, nx, ch]."""
        which = BUFFERS[name]
        if name in ("grad", "params"):
            out = np.empty(self.param_count(), dtype=np.float32)
            _check(lib().gng_net_read(self._h, which, _ptr(out), out.nbytes), "gng_net_read")
            return out
        kind, ch = _MESH[name]
        b = self.max_boxes if boxes is None else boxes
        n = self.grid_size
        raw = np.empty((b, n, n, n, ch), dtype=np.uint16 if kind == "bf16" else np.float32)
        _check(lib().gng_net_read(self._h, which, _ptr(raw), raw.nbytes), "gng_net_read")
        return bf16_to_f32(raw).reshape(raw.shape) if kind == "bf16" else raw

    def write(self, name: str, a: np.ndarray) -> None:
        """Write a mesh buffer from float32 values (rounded to bf16 where stored so)."""
        which = BUFFERS[name]
        kind, ch = _MESH[name]
        a = np.ascontiguousarray(a, dtype=np.float32)
        assert a.shape[-1] == ch and a.shape[1:4] == (self.grid_size,) * 3
        raw = f32_to_bf16(a).reshape(a.shape) if kind == "bf16" else a
        _check(lib().gng_net_write(self._h, which, _ptr(raw), raw.nbytes), "gng_net_write")

    def input_from_meshes(self, box: int, d_mass: int, d_phi: int, rho_mean: float, phi_scale: float,
                          stream: int = 0) -> None:
        _check(lib().gng_net_input(self._h, int(box), d_mass, d_phi, float(rho_mean), float(phi_scale),
                                   stream or None), "gng_net_input")

    def input_from_moments(self, box: int, d_mass: int, d_phi: int, rho_mean: float, phi_scale: float,
                           d_mom: int, stream: int = 0) -> None:
        """Input with the moment channels from per-cell moments (Plan.cell_moments_device)."""
        _check(lib().gng_net_input_moments(self._h, int(box), d_mass, d_phi, float(rho_mean), float(phi_scale),
                                           d_mom, stream or None), "gng_net_input_moments")

    def input_from_particles(self, box: int, d_mass: int, d_phi: int, rho_mean: float, phi_scale: float,
                             x: int, y: int, z: int, m: int, n: int, box_size: float, stream: int = 0) -> None:
        """Input with the particle-moment channels (gng_net_input_particles)."""
        _check(lib().gng_net_input_particles(self._h, int(box), d_mass, d_phi, float(rho_mean), float(phi_scale),
                                             x, y, z, m, int(n), float(box_size), stream or None),
               "gng_net_input_particles")

    def forward(self, boxes: int = 1, stream: int = 0) -> None:
        _check(lib().gng_net_forward(self._h, int(boxes), stream or None), "gng_net_forward")

    def infer(self, boxes: int = 1, stream: int = 0) -> None:
        """Forward pass without the pre-activations (inference)."""
        _check(lib().gng_net_infer(self._h, int(boxes), stream or None), "gng_net_infer")

    def loss_grad(self, pos4: int, tgt4: int, n: int, box_size: float, out_scale: float, stream: int = 0) -> None:
        _check(lib().gng_net_loss_grad(self._h, pos4, tgt4, int(n), float(box_size), float(out_scale),
                                       stream or None), "gng_net_loss_grad")

    def backward(self, boxes: int = 1, stream: int = 0) -> None:
        _check(lib().gng_net_backward(self._h, int(boxes), stream or None), "gng_net_backward")

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, stream: int = 0) -> None:
        _check(lib().gng_net_adam(self._h, float(lr), float(beta1), float(beta2), float(eps), stream or None),
               "gng_net_adam")

    def train_step(self, boxes: int, pos4: int, tgt4: int, n: int, box_size: float, out_scale: float, lr: float,
                   stream: int = 0) -> None:
        _check(lib().gng_net_train_step(self._h, int(boxes), pos4, tgt4, int(n), float(box_size),
                                        float(out_scale), float(lr), stream or None), "gng_net_train_step")

    def loss(self) -> float:
        v = ctypes.c_double()
        _check(lib().gng_net_loss(self._h, ctypes.byref(v)), "gng_net_loss")
        return float(v.value)

    def apply(self, box: int, x: int, y: int, z: int, n: int, box_size: float, out_scale: float, ax: int, ay: int,
              az: int, stream: int = 0) -> None:
        _check(lib().gng_net_apply(self._h, int(box), x, y, z, int(n), float(box_size), float(out_scale), ax, ay,
                                   az, stream or None), "gng_net_apply")

    def apply_field(self, box: int, x: int, y: int, z: int, n: int, box_size: float, out_scale: float, rho: int,
                    rho_max: float, ax: int, ay: int, az: int, stream: int = 0) -> None:
        """apply() for particles whose base cell holds at most rho_max (rho =
        device mass-per-cell mesh of the box); the others keep their force."""
        _check(lib().gng_net_apply_field(self._h, int(box), x, y, z, int(n), float(box_size), float(out_scale),
                                         rho or None, float(rho_max), ax, ay, az, stream or None),
               "gng_net_apply_field")

    def device_buffer(self, name: str) -> int:
        p = ctypes.c_void_p()
        _check(lib().gng_net_device_buffer(self._h, BUFFERS[name], ctypes.byref(p)), "gng_net_device_buffer")
        return int(p.value)


def _dcopy(dst: int, src: int, nbytes: int) -> None:
    _check(lib().gng_dev_copy(dst, src, int(nbytes), 2), "gng_dev_copy")


# ---------------------------------------------------------------------------
# data generation, training and evaluation
# ---------------------------------------------------------------------------

@dataclass
class Sample:
    """One box: particles on the device, PM and TreePM forces on the host."""
    n: int
    n_lattice: int
    box_size: float
    dev: List[DeviceArray]          # x, y, z, m (float32)
    f_pm: np.ndarray                # (n, 3) float32, plain PM
    f_true: np.ndarray              # (n, 3) float32, TreePM (long range + exact short range)
    pos: np.ndarray                 # (n, 3) float32
    in_field: np.ndarray            # (n,) bool, base cell below the dense threshold
    near_dense: np.ndarray          # (n,) bool, base cell is a dense cell or one of its 26 neighbours
    times: Dict[str, float] = field(default_factory=dict)


def force_errors(f: np.ndarray, f_true: np.ndarray, n_lattice: int, field=None, near=None) -> Dict[str, float]:
    """rms_rel = sqrt(sum|f - t|^2 / sum|t|^2) and the per-particle relative
    error |f - t| / |t| (median, mean) over all particles, the lattice and
    sphere members, and the field (particles in cells below the dense
    threshold) and dense populations when `field` (bool mask) is given; with
    `near` (base cell next to a dense cell) the field splits further into
    field_near_ and field_far_."""
    f = np.asarray(f, np.float64)
    t = np.asarray(f_true, np.float64)
    d = np.linalg.norm(f - t, axis=1)
    tn = np.maximum(np.linalg.norm(t, axis=1), 1e-30)
    r = d / tn

    def pop(sel, tag):
        if sel is None or not np.any(sel):
            return {}
        return {f"{tag}rms_rel": float(np.sqrt(np.sum(d[sel] ** 2) / np.sum(tn[sel] ** 2))),
                f"{tag}median_rel": float(np.median(r[sel])), f"{tag}mean_rel": float(np.mean(r[sel])),
                f"{tag}count": int(np.count_nonzero(sel))}

    out = pop(np.ones(len(r), bool), "")
    if 0 < n_lattice < len(r):
        lat = np.zeros(len(r), bool)
        lat[:n_lattice] = True
        out.update(pop(lat, "lattice_"))
        out.update(pop(~lat, "sphere_"))
    if field is not None:
        out.update(pop(field, "field_"))
        out.update(pop(~field, "dense_"))
        if near is not None:
            out.update(pop(field & near, "field_near_"))
            out.update(pop(field & ~near, "field_far_"))
    return out


class LearnedCorrection:
    """PM + learned sub-cell correction, trained toward TreePM forces.

    Boxes follow workloads.learned_box (one particle and one mesh cell per
    (Mpc/h)^3).  For every box the GPU computes the plain PM force (the
    method without the network, mesh = cells^3, CIC, no filter), the TreePM
    force (long range with r_s = r_split_cells cells, CIC-deconvolved, plus
    the exact-split short range summed within 5 r_s, Plummer softening eps)
    and the network input (log mass, asinh potential) from the same PM solve.
    Training minimises  mean_p w_p |F_PM + out_scale CIC(y) - F_TreePM|^2  with
    Adam, w_p = 1 / (|F_TreePM|^2 + f0^2) for field particles (base cell below
    (1 + dense_contrast) times the mean mass) and 0 in denser cells, where the
    residual depends on the arrangement inside the cell; f0 = out_scale = the
    median TreePM force of the field particles of the first training box.
    """

    def __init__(self, cells: int = 64, r_split_cells: float = 1.25, eps: float = 0.05, g: float = 1.0,
                 device: int = 0, batch: int = 8, dense_contrast: float = 8.0, phi_scale: float = 100.0):
        self.cells = int(cells)
        self.nm = self.cells
        self.L = float(cells)
        self.g = float(g)
        self.r_s = r_split_cells * self.L / self.nm
        self.r_cut = 5.0 * self.r_s
        self.eps2 = float(eps) ** 2
        self.device = device
        self.batch = int(batch)
        self.pm = Plan(self.nm, self.L, device=device, keep_density=True, keep_potential=True)
        self.lr = Plan(self.nm, self.L, r_split=self.r_s, deconvolve=True, device=device)
        self.sr = ShortRange(device=device)
        self.net = Net(self.nm, self.batch, device=device)
        # per-cell particle moments of the last PM solve (5 doubles per cell)
        self._mom = DeviceArray((self.nm ** 3 * 5,), np.float64)
        self.f0: Optional[float] = None
        # potential unit: phi_scale * G * (mean mass per cell) / (cell size)
        self.phi_scale = float(phi_scale) * self.g
        # particles whose base cell holds more than (1 + dense_contrast) times
        # the mean mass are left out of the loss: their residual is set by the
        # arrangement inside the cell, which the mesh does not resolve
        self.dense_contrast = float(dense_contrast)
        self._groups = []

    # -- per-box forces -----------------------------------------------------
    def _forces(self, x: DeviceArray, y: DeviceArray, z: DeviceArray, m: DeviceArray, n: int, net: Net,
                slot: int, rho_mean: float):
        a_pm = [DeviceArray((n,), np.float32) for _ in range(3)]
        a_tp = [DeviceArray((n,), np.float32) for _ in range(3)]
        tmp = [DeviceArray((n,), np.float32) for _ in range(3)]
        t0 = time.perf_counter()
        self.pm.solve_device(x.ptr, y.ptr, z.ptr, m.ptr, n, a_pm[0].ptr, a_pm[1].ptr, a_pm[2].ptr, self.g)
        dm, dp = self.pm.device_meshes()
        self.pm.cell_moments_device(self._mom.ptr)
        net.input_from_moments(slot, dm, dp, rho_mean, self.phi_scale, self._mom.ptr)
        device_synchronize()
        t1 = time.perf_counter()
        self.lr.solve_device(x.ptr, y.ptr, z.ptr, m.ptr, n, a_tp[0].ptr, a_tp[1].ptr, a_tp[2].ptr, self.g)
        self.sr.accelerations_device(x.ptr, y.ptr, z.ptr, m.ptr, n, self.L, self.r_s, self.r_cut, self.eps2, self.g,
                                     SR_EXACT, tmp[0].ptr, tmp[1].ptr, tmp[2].ptr)
        _axpy_device([a.ptr for a in a_tp], tmp, n)
        device_synchronize()
        t2 = time.perf_counter()
        f_pm = np.stack([a.numpy() for a in a_pm], axis=1)
        f_tp = np.stack([a.numpy() for a in a_tp], axis=1)
        return f_pm, f_tp, {"pm_s": t1 - t0, "treepm_s": t2 - t1, "sr_pairs": self.sr.last_pair_count()}

    def sample(self, seed: int, cells: Optional[int] = None, net: Optional[Net] = None, slot: int = 0,
               workers: int = -1) -> Sample:
        cells = self.cells if cells is None else cells
        if cells != self.cells:
            raise ValueError("sample() uses the solver's own box size; build another LearnedCorrection")
        p, n_lat = workloads.learned_box(seed, cells=cells, workers=workers)
        dev = [DeviceArray.from_host(a) for a in (p.x, p.y, p.z, p.m)]
        rho_mean = float(np.sum(p.m, dtype=np.float64)) / self.nm ** 3
        f_pm, f_tp, times = self._forces(*dev, p.n, self.net if net is None else net, slot, rho_mean)
        pos = np.stack([p.x, p.y, p.z], axis=1)
        mass = self.pm.density()
        c = np.floor(pos.astype(np.float64) * (self.nm / self.L)).astype(np.int64) % self.nm
        dense = mass > (1.0 + self.dense_contrast) * rho_mean
        fld = ~dense[c[:, 2], c[:, 1], c[:, 0]]
        halo = dense.copy()
        for ax in range(3):
            halo = halo | np.roll(halo, 1, axis=ax) | np.roll(halo, -1, axis=ax)
        near = halo[c[:, 2], c[:, 1], c[:, 0]]
        return Sample(p.n, n_lat, self.L, dev, f_pm, f_tp, pos, fld, near, times)

    def _records(self, s: Sample, slot: int):
        """Training records sorted by base cell: pos4 (x, y, z, slot), tgt4
        (residual F_true - F_PM, weight 1 / (|F_true|^2 + f0^2))."""
        c = np.floor(s.pos.astype(np.float64) * (self.nm / self.L)).astype(np.int64) % self.nm
        key = (c[:, 2] * self.nm + c[:, 1]) * self.nm + c[:, 0]
        order = np.argsort(key, kind="stable")
        pos4 = np.empty((s.n, 4), np.float32)
        pos4[:, :3] = s.pos[order]
        pos4[:, 3] = slot
        tgt4 = np.empty((s.n, 4), np.float32)
        tgt4[:, :3] = (s.f_true - s.f_pm)[order]
        tn2 = np.sum(s.f_true.astype(np.float64) ** 2, axis=1)[order]
        tgt4[:, 3] = np.where(s.in_field[order], 1.0 / (tn2 + self.f0 ** 2), 0.0)
        return pos4, tgt4

    def build_training_set(self, seeds, workers: int = -1, log=None) -> Dict[str, float]:
        """Generate the training boxes in groups of `batch` (inputs staged on the device)."""
        seeds = list(seeds)
        self._groups = []
        M = self.nm ** 3
        t_gen = t_gpu = 0.0
        pairs = 0
        for g0 in range(0, len(seeds), self.batch):
            grp = seeds[g0:g0 + self.batch]
            pos_l, tgt_l = [], []
            for j, sd in enumerate(grp):
                t = time.perf_counter()
                s = self.sample(sd, slot=j, workers=workers)
                t_gen += time.perf_counter() - t - s.times["pm_s"] - s.times["treepm_s"]
                t_gpu += s.times["pm_s"] + s.times["treepm_s"]
                pairs += s.times["sr_pairs"]
                if self.f0 is None:
                    self.f0 = float(np.median(np.linalg.norm(s.f_true[s.in_field].astype(np.float64), axis=1)))
                p4, t4 = self._records(s, j)
                pos_l.append(p4)
                tgt_l.append(t4)
            x0 = DeviceArray((len(grp) * M * X0_CHANNELS,), np.uint16)
            _dcopy(x0.ptr, self.net.device_buffer("x0"), len(grp) * M * X0_CHANNELS * 2)
            pos4 = DeviceArray.from_host(np.concatenate(pos_l))
            tgt4 = DeviceArray.from_host(np.concatenate(tgt_l))
            self._groups.append((len(grp), x0, pos4, tgt4, pos4.shape[0]))
            if log:
                log(f"training group {len(self._groups)}: seeds {grp[0]}..{grp[-1]}")
        return {"boxes": len(seeds), "gen_cpu_s": t_gen, "forces_gpu_s": t_gpu, "sr_pairs": pairs}

    def train(self, steps: int, lr: float = 1e-3, params=None, seed: int = 1234, log=None,
              log_every: int = 100) -> List[float]:
        """Adam over the staged groups, one group per step (cyclic).  Returns
        the loss of every logged step (read back only at log points)."""
        if not self._groups:
            raise RuntimeError("build_training_set() first")
        self.net.set_params(init_params(seed) if params is None else params)
        M = self.nm ** 3
        hist = []
        for it in range(steps):
            nb, x0, pos4, tgt4, n = self._groups[it % len(self._groups)]
            _dcopy(self.net.device_buffer("x0"), x0.ptr, nb * M * X0_CHANNELS * 2)
            cur_lr = lr * (0.1 if it >= int(0.8 * steps) else 1.0)
            self.net.train_step(nb, pos4.ptr, tgt4.ptr, n, self.L, self.f0, cur_lr)
            if log_every and (it % log_every == 0 or it == steps - 1):
                hist.append(self.net.loss())
                if log:
                    log(f"step {it}: loss {hist[-1]:.5f}")
        device_synchronize()
        return hist

    def training_loss(self) -> float:
        """Mean loss of the current weights over every staged group (no update)."""
        M = self.nm ** 3
        tot = cnt = 0.0
        for nb, x0, pos4, tgt4, n in self._groups:
            _dcopy(self.net.device_buffer("x0"), x0.ptr, nb * M * X0_CHANNELS * 2)
            self.net.forward(nb)
            self.net.loss_grad(pos4.ptr, tgt4.ptr, n, self.L, self.f0)
            tot += self.net.loss() * n
            cnt += n
        return tot / cnt

    def evaluate(self, seed: int, cells: Optional[int] = None, workers: int = -1) -> Dict[str, float]:
        """Held-out box: errors of PM and of PM + network against TreePM, and
        the GPU time of each method (host wall clock around synchronised calls)."""
        cells = self.cells if cells is None else int(cells)
        other = self if cells == self.cells else LearnedCorrection(
            cells=cells, r_split_cells=self.r_s * self.nm / self.L, eps=math.sqrt(self.eps2), g=self.g,
            device=self.device, batch=1)
        net = other.net
        if other is not self:
            net.set_params(self.net.params())
            other.f0 = self.f0
            other.phi_scale = self.phi_scale
            other.dense_contrast = self.dense_contrast
        s = other.sample(seed, net=net, slot=0, workers=workers)
        n = s.n
        corr = [DeviceArray.from_host(np.ascontiguousarray(s.f_pm[:, a])) for a in range(3)]
        x, y, z, m = s.dev
        rho_mean = float(np.sum(m.numpy(), dtype=np.float64)) / cells ** 3
        rho_max = (1.0 + self.dense_contrast) * rho_mean
        dm, _ = other.pm.device_meshes()
        net.infer(1)
        # the correction goes to the field only: dense-cell particles carry no
        # weight in the loss and keep the PM force
        net.apply_field(0, x.ptr, y.ptr, z.ptr, n, other.L, self.f0, dm, rho_max, corr[0].ptr, corr[1].ptr,
                        corr[2].ptr)
        device_synchronize()
        f_corr = np.stack([c.numpy() for c in corr], axis=1)
        # time the pieces once more, warm: PM, network input, network + apply
        a = [DeviceArray((n,), np.float32) for _ in range(3)]
        device_synchronize()
        t0 = time.perf_counter()
        other.pm.solve_device(x.ptr, y.ptr, z.ptr, m.ptr, n, a[0].ptr, a[1].ptr, a[2].ptr, self.g)
        device_synchronize()
        t_pm = time.perf_counter() - t0
        t0 = time.perf_counter()
        dm, dp = other.pm.device_meshes()
        other.pm.cell_moments_device(other._mom.ptr)
        net.input_from_moments(0, dm, dp, rho_mean, other.phi_scale, other._mom.ptr)
        device_synchronize()
        t_in = time.perf_counter() - t0
        t0 = time.perf_counter()
        net.infer(1)
        net.apply_field(0, x.ptr, y.ptr, z.ptr, n, other.L, self.f0, dm, rho_max, a[0].ptr, a[1].ptr, a[2].ptr)
        device_synchronize()
        t_net = time.perf_counter() - t0
        t0 = time.perf_counter()
        other.lr.solve_device(x.ptr, y.ptr, z.ptr, m.ptr, n, a[0].ptr, a[1].ptr, a[2].ptr, self.g)
        device_synchronize()
        t_lr = time.perf_counter() - t0
        t0 = time.perf_counter()
        other.sr.accelerations_device(x.ptr, y.ptr, z.ptr, m.ptr, n, other.L, other.r_s, other.r_cut, other.eps2,
                                      self.g, SR_EXACT, a[0].ptr, a[1].ptr, a[2].ptr)
        device_synchronize()
        t_sr = time.perf_counter() - t0
        return {"seed": seed, "cells": cells, "n": n,
                "pm": force_errors(s.f_pm, s.f_true, s.n_lattice, s.in_field, s.near_dense),
                "pm_net": force_errors(f_corr, s.f_true, s.n_lattice, s.in_field, s.near_dense),
                "time_pm_s": t_pm, "time_input_s": t_in, "time_net_s": t_net, "time_treepm_long_s": t_lr,
                "time_short_range_s": t_sr,
                "sr_pairs": other.sr.last_pair_count()}
