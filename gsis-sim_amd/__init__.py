"""Python binding of the MI355X kinetic step (ctypes over include/gsis_kinetic.h).

The library is libgsis_kinetic.so in this directory (built by `make -C gsis-sim_amd`). Nothing
here computes physics: every call goes to the C ABI, whose backend runs the HIP kernels. Loading
fails loudly when the library is missing; backend and solver creation fail loudly when no GPU is
present (there is no CPU fallback in this package).

Mirrors (reference tree):
  Setup   -> build_kinetic_setup (src/kinetic/kinetic_setup.hpp:118-120)
  Backend -> KineticBackend method set (src/kinetic/kinetic_backend.hpp:59-100)
  Solver  -> KineticSolver (src/kinetic/kinetic_solver.hpp:61-115)
"""

from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgsis_kinetic.so")

N_MOMENTS = 15
N_BSUMS = 20


class GskError(RuntimeError):
    """A failed library call (the reference raises gsis::Error)."""


class CsrEntry(C.Structure):
    _fields_ = [("other", C.c_int32), ("bface", C.c_int32), ("s", C.c_double * 3), ("ds", C.c_double * 3),
                ("dother", C.c_double * 3), ("c", C.c_double * 3)]


class BFace(C.Structure):
    _fields_ = [("owner", C.c_int32), ("kind", C.c_int32), ("mirror_xor", C.c_int32), ("patch", C.c_int32),
                ("n", C.c_double * 3), ("area", C.c_double), ("d", C.c_double * 3)]


class InflowState(C.Structure):
    _fields_ = [("scale", C.c_double), ("rho", C.c_double), ("ux", C.c_double), ("uy", C.c_double),
                ("uz", C.c_double), ("T", C.c_double)]


class Timings(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("moments", "boundary", "gradient", "residual", "sweep_forward", "sweep_backward", "other")]


class DeqCounts(C.Structure):
    _fields_ = [("corrected", C.c_longlong), ("failed", C.c_longlong), ("max_iterations", C.c_int)]


class SetupView(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int32), ("nv", C.c_int32), ("n_bfaces", C.c_int32), ("n_blocks", C.c_int32),
        ("orbit_size", C.c_int32), ("velocity_chunk", C.c_int64),
        ("xi_x", C.POINTER(C.c_double)), ("xi_y", C.POINTER(C.c_double)), ("xi_z", C.POINTER(C.c_double)),
        ("weight", C.c_double),
        ("csr_offsets", C.POINTER(C.c_int32)), ("csr", C.POINTER(CsrEntry)), ("eps2", C.POINTER(C.c_double)),
        ("volume", C.POINTER(C.c_double)), ("vdt", C.POINTER(C.c_double)), ("bfaces", C.POINTER(BFace)),
        ("block_offsets", C.POINTER(C.c_int32)), ("n_sweep_dirs", C.c_int32),
        ("sweep_order", C.POINTER(C.c_int32)), ("sweep_pos", C.POINTER(C.c_int32)),
        ("sweep_dir_of", C.POINTER(C.c_uint8)), ("grid_to_storage", C.POINTER(C.c_int32)),
        ("grid_n", C.c_int32 * 3), ("n_maxwell_faces", C.c_int32),
        ("maxwell_offset", C.POINTER(C.c_int64)), ("maxwell_specular", C.POINTER(C.c_double)),
        ("deq_lo", C.c_double * 3), ("deq_hi", C.c_double * 3), ("deq_d", C.c_double * 3),
        ("deq_criterion", C.c_double), ("deq_max_iterations", C.c_int32), ("deq_tolerance", C.c_double),
        ("discrete_eq", C.c_int32), ("deq_fallback_collisionless", C.c_int32), ("limiter_mode", C.c_int32),
        ("positivity", C.c_int32), ("diag_scale", C.c_double), ("interface_damping", C.c_double),
        ("cold_bgk_temperature", C.c_double), ("mu0", C.c_double), ("omega", C.c_double), ("Pr", C.c_double),
        ("block_passes", C.c_int32), ("negative_check", C.c_int32), ("precision_fp64", C.c_int32),
        ("moments_mfma", C.c_int32),
        ("weight_rel", C.POINTER(C.c_double)), ("grid_nodes", C.POINTER(C.c_double)),
    ]


class IterationInfo(C.Structure):
    _fields_ = [
        ("iteration", C.c_longlong), ("epsilon", C.c_double), ("eps_rho", C.c_double), ("eps_u", C.c_double),
        ("eps_T", C.c_double), ("has_epsilon", C.c_int32), ("residual", C.c_double), ("has_residual", C.c_int32),
        ("total_mass", C.c_double), ("mass_scale", C.c_double), ("q_mean", C.c_double * 3),
        ("sigma_mean", C.c_double * 6), ("seconds", C.c_double), ("timings", Timings),
        ("deq_corrected", C.c_longlong), ("deq_failed", C.c_longlong), ("deq_max_iterations", C.c_int32),
    ]


_FN = C.CFUNCTYPE


class BackendOps(C.Structure):
    _fields_ = [
        ("ctx", C.c_void_p),
        ("name", C.c_void_p), ("chunk_width", C.c_void_p), ("memory_bytes", C.c_void_p),
        ("set_equilibrium", C.c_void_p), ("set_distribution", C.c_void_p), ("get_distribution", C.c_void_p),
        ("compute_moments", C.c_void_p), ("scale_distribution", C.c_void_p), ("update_cell_params", C.c_void_p),
        ("set_inflow", C.c_void_p), ("boundary_sums", C.c_void_p), ("sweep", C.c_void_p),
        ("gsis_correction", C.c_void_p), ("internal_face_mass_flux", C.c_void_p), ("take_timings", C.c_void_p),
        ("deq_counts", C.c_void_p), ("last_error", C.c_void_p),
    ]


# exported symbols of include/gsis_kinetic.h (checked by the CPU test suite)
EXPORTS = [
    "gsk_last_error", "gsk_version", "gsk_setup_from_case", "gsk_setup_destroy", "gsk_setup_view_get",
    "gsk_setup_sizes", "gsk_setup_cell_new_to_old", "gsk_setup_storage_to_grid", "gsk_setup_cell_centers",
    "gsk_setup_mass_flow_scale", "gsk_setup_write_pmesh", "gsk_setup_describe", "gsk_setup_bface_geometry",
    "gsk_backend_create", "gsk_backend_destroy",
    "gsk_backend_name", "gsk_backend_chunk_width", "gsk_backend_memory_bytes", "gsk_backend_set_equilibrium",
    "gsk_backend_set_distribution", "gsk_backend_get_distribution", "gsk_backend_compute_moments",
    "gsk_backend_scale_distribution", "gsk_backend_update_cell_params", "gsk_backend_set_inflow",
    "gsk_backend_boundary_sums", "gsk_backend_sweep", "gsk_backend_gsis_correction",
    "gsk_backend_internal_face_mass_flux", "gsk_backend_take_timings", "gsk_backend_deq_counts",
    "gsk_backend_gsis_correction_noalloc",
    "gsk_solver_create", "gsk_solver_create_with_backend", "gsk_solver_destroy", "gsk_solver_initialize_uniform",
    "gsk_solver_initialize_equilibrium", "gsk_solver_prepare", "gsk_solver_update_mass_flow_pressure",
    "gsk_solver_sweep", "gsk_solver_patch_mass_flow", "gsk_solver_moments", "gsk_solver_boundary_fluxes",
    "gsk_solver_apply_gsis_correction", "gsk_solver_distribution", "gsk_solver_set_distribution",
    "gsk_solver_residual_per_cell", "gsk_solver_section_mass_flow", "gsk_solver_section_cuts", "gsk_solver_backend", "gsk_solver_element_size", "gsk_solver_distribution_device",
    "gsk_solver_set_distribution_device", "gsk_solver_residual_device", "gsk_solver_precondition_device",
    "gsk_solver_mu0", "gsk_solver_set_mu0", "gsk_solver_set_accommodation", "gsk_solver_set_newton_tolerances",
    "gsk_solver_set_boundary_iterations", "gsk_solver_set_device_prepare", "gsk_solver_device_prepare",
    "gsk_solver_freeze_limiters", "gsk_solver_adjoint_prepare",
    "gsk_solver_adjoint_apply_device",
]

# fixed device scratch of gsk_backend_gsis_correction_noalloc (GSK_NOALLOC_SCRATCH_BYTES, GSK_NOALLOC_MAX_CELLS)
NOALLOC_SCRATCH_BYTES = 16 * 1024 * 1024
NOALLOC_MAX_CELLS = (NOALLOC_SCRATCH_BYTES - 256) // 80

_lib_handle = None


def lib() -> C.CDLL:
    """Load libgsis_kinetic.so (raises when it has not been built)."""
    global _lib_handle
    if _lib_handle is not None:
        return _lib_handle
    if not os.path.exists(LIB_PATH):
        raise GskError(f"{LIB_PATH} is missing: build it with `make -C {_HERE}` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    L.gsk_last_error.restype = C.c_char_p
    L.gsk_version.restype = C.c_char_p
    L.gsk_backend_name.restype = C.c_char_p
    L.gsk_solver_backend.restype = C.c_void_p
    L.gsk_solver_backend.argtypes = [C.c_void_p]
    _lib_handle = L
    return L


def _check(status: int) -> None:
    if status != 0:
        raise GskError(lib().gsk_last_error().decode(errors="replace"))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


class Setup:
    """Kinetic setup from a case TOML + .pmesh (overrides: "key = value" lines for [kinetic])."""

    def __init__(self, toml_path: str, overrides: str | dict | None = None):
        if isinstance(overrides, dict):
            overrides = "\n".join(f"{k} = {_toml_value(v)}" for k, v in overrides.items())
        h = C.c_void_p()
        _check(lib().gsk_setup_from_case(str(toml_path).encode(), overrides.encode() if overrides else None,
                                         C.byref(h)))
        self._h = h
        nc, nv, nb, nblk, npat = (C.c_int32() for _ in range(5))
        _check(lib().gsk_setup_sizes(self._h, C.byref(nc), C.byref(nv), C.byref(nb), C.byref(nblk), C.byref(npat)))
        self.n_cells, self.nv, self.n_bfaces, self.n_blocks, self.n_patches = (
            nc.value, nv.value, nb.value, nblk.value, npat.value)

    def view(self) -> SetupView:
        v = SetupView()
        _check(lib().gsk_setup_view_get(self._h, C.byref(v)))
        return v

    def cell_new_to_old(self) -> np.ndarray:
        out = np.empty(self.n_cells, np.int32)
        _check(lib().gsk_setup_cell_new_to_old(self._h, _iptr(out)))
        return out

    def storage_to_grid(self) -> np.ndarray:
        out = np.empty(self.nv, np.int32)
        _check(lib().gsk_setup_storage_to_grid(self._h, _iptr(out)))
        return out

    def cell_centers(self) -> np.ndarray:
        out = np.empty((self.n_cells, 3), np.float64)
        _check(lib().gsk_setup_cell_centers(self._h, _dptr(out)))
        return out

    def mass_flow_factor(self) -> float:
        out = C.c_double()
        _check(lib().gsk_setup_mass_flow_scale(self._h, C.byref(out)))
        return out.value

    def describe(self) -> dict:
        """Case name and files, sizes, Kn / delta, reference scales, [gsis] settings, [kinetic]
        parameters and patches (gsk_setup_describe)."""
        need = C.c_size_t()
        _check(lib().gsk_setup_describe(self._h, None, C.c_size_t(0), C.byref(need)))
        buf = C.create_string_buffer(need.value)
        _check(lib().gsk_setup_describe(self._h, buf, C.c_size_t(need.value), C.byref(need)))
        return json.loads(buf.value.decode())

    def bface_geometry(self):
        """(patch index, area in mesh units^2, outward unit normal) per boundary face, original numbering."""
        nb = self.n_bfaces
        patch = np.empty(nb, np.int32)
        area = np.empty(nb, np.float64)
        normal = np.empty((nb, 3), np.float64)
        _check(lib().gsk_setup_bface_geometry(self._h, _iptr(patch), _dptr(area), _dptr(normal)))
        return patch, area, normal

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib_handle is not None:
            _lib_handle.gsk_setup_destroy(h)
            self._h = None


def _toml_value(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, str):
        return f'"{v}"'
    return repr(v)


class Backend:
    """HIP backend with the reference KineticBackend method set (host data: float64 numpy)."""

    def __init__(self, setup: Setup, _handle=None, _owner=None):
        self.setup = setup
        self._owner = _owner
        if _handle is None:
            h = C.c_void_p()
            v = setup.view()
            _check(lib().gsk_backend_create(C.byref(v), C.byref(h)))
            self._h = h
            self._owned = True
        else:
            self._h = C.c_void_p(_handle)
            self._owned = False

    @property
    def name(self) -> str:
        return lib().gsk_backend_name(self._h).decode()

    def chunk_width(self) -> int:
        k = C.c_int32()
        _check(lib().gsk_backend_chunk_width(self._h, C.byref(k)))
        return k.value

    def memory_bytes(self) -> int:
        m = C.c_uint64()
        _check(lib().gsk_backend_memory_bytes(self._h, C.byref(m)))
        return m.value

    def set_equilibrium(self, states: np.ndarray) -> None:
        s = np.ascontiguousarray(states, np.float64).ravel()
        _check(lib().gsk_backend_set_equilibrium(self._h, _dptr(s), C.c_size_t(s.size)))

    def set_distribution(self, f: np.ndarray) -> None:
        a = np.ascontiguousarray(f, np.float64).ravel()
        _check(lib().gsk_backend_set_distribution(self._h, _dptr(a), C.c_size_t(a.size)))

    def get_distribution(self) -> np.ndarray:
        out = np.empty(self.setup.n_cells * self.setup.nv, np.float64)
        _check(lib().gsk_backend_get_distribution(self._h, _dptr(out), C.c_size_t(out.size)))
        return out.reshape(self.setup.n_cells, self.setup.nv)

    def compute_moments(self) -> np.ndarray:
        out = np.empty(self.setup.n_cells * N_MOMENTS, np.float64)
        _check(lib().gsk_backend_compute_moments(self._h, _dptr(out), C.c_size_t(out.size)))
        return out.reshape(self.setup.n_cells, N_MOMENTS)

    def scale_distribution(self, s: float) -> None:
        _check(lib().gsk_backend_scale_distribution(self._h, C.c_double(s)))

    def update_cell_params(self) -> None:
        _check(lib().gsk_backend_update_cell_params(self._h))

    def set_inflow(self, inflow: np.ndarray) -> None:
        """inflow: (n_bfaces, 6) = scale, rho, ux, uy, uz, T."""
        a = np.ascontiguousarray(inflow, np.float64).reshape(-1, 6)
        _check(lib().gsk_backend_set_inflow(self._h, a.ctypes.data_as(C.POINTER(InflowState)), C.c_size_t(len(a))))

    def boundary_sums(self) -> np.ndarray:
        out = np.empty(self.setup.n_bfaces * N_BSUMS, np.float64)
        _check(lib().gsk_backend_boundary_sums(self._h, _dptr(out), C.c_size_t(out.size)))
        return out.reshape(self.setup.n_bfaces, N_BSUMS)

    def sweep(self) -> np.ndarray:
        out = np.empty(self.setup.n_cells, np.float64)
        _check(lib().gsk_backend_sweep(self._h, _dptr(out), C.c_size_t(out.size)))
        return out

    def gsis_correction(self, w_half: np.ndarray, w_new: np.ndarray):
        wh = np.ascontiguousarray(w_half, np.float64).ravel()
        wn = np.ascontiguousarray(w_new, np.float64).ravel()
        flags = np.zeros(self.setup.n_cells, np.int32)
        rep = C.c_longlong()
        _check(lib().gsk_backend_gsis_correction(self._h, _dptr(wh), _dptr(wn), C.c_size_t(wh.size), _iptr(flags),
                                                 C.c_size_t(flags.size), C.byref(rep)))
        return rep.value, flags

    def gsis_correction_noalloc(self, w_half: np.ndarray, w_new: np.ndarray):
        """gsis_correction through the entry point that allocates no device memory (fixed per-device
        scratch; at most NOALLOC_MAX_CELLS cells)."""
        wh = np.ascontiguousarray(w_half, np.float64).ravel()
        wn = np.ascontiguousarray(w_new, np.float64).ravel()
        flags = np.zeros(self.setup.n_cells, np.int32)
        rep = C.c_longlong()
        _check(lib().gsk_backend_gsis_correction_noalloc(self._h, _dptr(wh), _dptr(wn), C.c_size_t(wh.size),
                                                         _iptr(flags), C.c_size_t(flags.size), C.byref(rep)))
        return rep.value, flags

    def internal_face_mass_flux(self, cells: np.ndarray, entries: np.ndarray) -> np.ndarray:
        c = np.ascontiguousarray(cells, np.int32)
        e = np.ascontiguousarray(entries, np.int32)
        out = np.empty(c.size, np.float64)
        _check(lib().gsk_backend_internal_face_mass_flux(self._h, _iptr(c), _iptr(e), C.c_size_t(c.size), _dptr(out)))
        return out

    def take_timings(self) -> dict:
        t = Timings()
        _check(lib().gsk_backend_take_timings(self._h, C.byref(t)))
        return {n: getattr(t, n) for n, _ in Timings._fields_}

    def deq_counts(self) -> tuple:
        d = DeqCounts()
        _check(lib().gsk_backend_deq_counts(self._h, C.byref(d)))
        return d.corrected, d.failed, d.max_iterations

    def close(self) -> None:
        if getattr(self, "_owned", False) and getattr(self, "_h", None) is not None and _lib_handle is not None:
            _lib_handle.gsk_backend_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()


def _info_dict(i: IterationInfo) -> dict:
    d = {n: getattr(i, n) for n, _ in IterationInfo._fields_ if n not in ("q_mean", "sigma_mean", "timings")}
    d["q_mean"] = list(i.q_mean)
    d["sigma_mean"] = list(i.sigma_mean)
    d["timings"] = {n: getattr(i.timings, n) for n, _ in Timings._fields_}
    return d


class Solver:
    """KineticSolver over the HIP backend, or over an external backend ops table (`ops`)."""

    def __init__(self, setup: Setup, ops: BackendOps | None = None):
        self.setup = setup
        h = C.c_void_p()
        if ops is None:
            _check(lib().gsk_solver_create(setup._h, C.byref(h)))
        else:
            self._ops = ops
            _check(lib().gsk_solver_create_with_backend(setup._h, C.byref(ops), C.byref(h)))
        self._h = h

    def backend(self) -> Backend:
        return Backend(self.setup, _handle=lib().gsk_solver_backend(self._h), _owner=self)

    def initialize_uniform(self, rho=1.0, u=(0.0, 0.0, 0.0), T=1.0) -> None:
        _check(lib().gsk_solver_initialize_uniform(self._h, C.c_double(rho), C.c_double(u[0]), C.c_double(u[1]),
                                                   C.c_double(u[2]), C.c_double(T)))

    def initialize_equilibrium(self, states: np.ndarray) -> None:
        s = np.ascontiguousarray(states, np.float64).ravel()
        _check(lib().gsk_solver_initialize_equilibrium(self._h, _dptr(s), C.c_size_t(s.size)))

    def prepare(self) -> dict:
        i = IterationInfo()
        _check(lib().gsk_solver_prepare(self._h, C.byref(i)))
        return _info_dict(i)

    def update_mass_flow_pressure(self) -> None:
        _check(lib().gsk_solver_update_mass_flow_pressure(self._h))

    def sweep(self, update_pressure: bool = True) -> dict:
        i = IterationInfo()
        _check(lib().gsk_solver_sweep(self._h, C.c_int32(1 if update_pressure else 0), C.byref(i)))
        return _info_dict(i)

    def patch_mass_flow(self) -> np.ndarray:
        out = np.empty(self.setup.n_patches, np.float64)
        _check(lib().gsk_solver_patch_mass_flow(self._h, _dptr(out), C.c_size_t(out.size)))
        return out

    def moments(self) -> np.ndarray:
        out = np.empty(self.setup.n_cells * N_MOMENTS, np.float64)
        _check(lib().gsk_solver_moments(self._h, _dptr(out), C.c_size_t(out.size)))
        return out.reshape(self.setup.n_cells, N_MOMENTS)

    def boundary_fluxes(self):
        nb = self.setup.n_bfaces
        fp, fm, fs = (np.empty(nb * 5, np.float64) for _ in range(3))
        _check(lib().gsk_solver_boundary_fluxes(self._h, _dptr(fp), _dptr(fm), _dptr(fs), C.c_size_t(nb)))
        return fp.reshape(nb, 5), fm.reshape(nb, 5), fs.reshape(nb, 5)

    def apply_gsis_correction(self, w_half: np.ndarray, w_new: np.ndarray) -> int:
        wh = np.ascontiguousarray(w_half, np.float64).ravel()
        wn = np.ascontiguousarray(w_new, np.float64).ravel()
        rep = C.c_longlong()
        _check(lib().gsk_solver_apply_gsis_correction(self._h, _dptr(wh), _dptr(wn), C.c_size_t(wh.size), C.byref(rep)))
        return rep.value

    def distribution(self) -> np.ndarray:
        out = np.empty(self.setup.n_cells * self.setup.nv, np.float64)
        _check(lib().gsk_solver_distribution(self._h, _dptr(out), C.c_size_t(out.size)))
        return out.reshape(self.setup.n_cells, self.setup.nv)

    def set_distribution(self, f: np.ndarray) -> None:
        a = np.ascontiguousarray(f, np.float64).ravel()
        _check(lib().gsk_solver_set_distribution(self._h, _dptr(a), C.c_size_t(a.size)))

    def residual_per_cell(self) -> np.ndarray:
        out = np.empty(self.setup.n_cells, np.float64)
        _check(lib().gsk_solver_residual_per_cell(self._h, _dptr(out), C.c_size_t(out.size)))
        return out

    def section_mass_flow(self, positions, axis: int = 0) -> np.ndarray:
        """Nondimensional mass flow along +axis through the sections at `positions` (L_ref units)."""
        pos = np.ascontiguousarray(positions, np.float64)
        out = np.empty(pos.size, np.float64)
        _check(lib().gsk_solver_section_mass_flow(self._h, C.c_int32(axis), _dptr(pos), C.c_size_t(pos.size),
                                                  _dptr(out)))
        return out

    def section_cuts(self, positions, axis: int = 0) -> dict:
        """Per section: face count, area, projected area (L_ref^2 units), area-weighted face position."""
        pos = np.ascontiguousarray(positions, np.float64)
        n = np.empty(pos.size, np.int64)
        area, proj, fpos = (np.empty(pos.size, np.float64) for _ in range(3))
        _check(lib().gsk_solver_section_cuts(self._h, C.c_int32(axis), _dptr(pos), C.c_size_t(pos.size),
                                             n.ctypes.data_as(C.POINTER(C.c_int64)), _dptr(area), _dptr(proj),
                                             _dptr(fpos)))
        return dict(n_faces=n, area=area, projected_area=proj, face_position=fpos)

    # ---- device-resident operators (device pointers, n = cells x velocities values of element_size) ----
    def element_size(self) -> int:
        out = C.c_int32()
        _check(lib().gsk_solver_element_size(self._h, C.byref(out)))
        return out.value

    def distribution_device(self, ptr: int, n: int) -> None:
        _check(lib().gsk_solver_distribution_device(self._h, C.c_void_p(ptr), C.c_size_t(n)))

    def set_distribution_device(self, ptr: int, n: int) -> None:
        _check(lib().gsk_solver_set_distribution_device(self._h, C.c_void_p(ptr), C.c_size_t(n)))

    def residual_device(self, ptr: int, n: int) -> None:
        _check(lib().gsk_solver_residual_device(self._h, C.c_void_p(ptr), C.c_size_t(n)))

    def precondition_device(self, r_ptr: int, out_ptr: int, n: int) -> None:
        _check(lib().gsk_solver_precondition_device(self._h, C.c_void_p(r_ptr), C.c_void_p(out_ptr), C.c_size_t(n)))

    def mu0(self) -> float:
        out = C.c_double()
        _check(lib().gsk_solver_mu0(self._h, C.byref(out)))
        return out.value

    def set_mu0(self, mu0: float) -> None:
        _check(lib().gsk_solver_set_mu0(self._h, C.c_double(mu0)))

    def set_accommodation(self, patch: int, alpha: float) -> None:
        _check(lib().gsk_solver_set_accommodation(self._h, C.c_int32(patch), C.c_double(alpha)))

    def set_newton_tolerances(self, cell_tol: float, inflow_tol: float) -> None:
        _check(lib().gsk_solver_set_newton_tolerances(self._h, C.c_double(cell_tol), C.c_double(inflow_tol)))

    def set_boundary_iterations(self, n: int) -> None:
        _check(lib().gsk_solver_set_boundary_iterations(self._h, C.c_int64(n)))

    def set_device_prepare(self, on: bool) -> None:
        """Run each iteration's prepare on the backend's device (on, the default where it offers one)
        or on the host."""
        _check(lib().gsk_solver_set_device_prepare(self._h, C.c_int32(1 if on else 0)))

    def prepared_on_device(self) -> bool:
        """Whether the last prepare ran on the backend's device."""
        out = C.c_int32(0)
        _check(lib().gsk_solver_device_prepare(self._h, C.byref(out)))
        return bool(out.value)

    def freeze_limiters(self, on: bool) -> None:
        """Record every limiter branch at the next residual_device and follow it afterwards (on), or
        restore the normal limiter (off)."""
        _check(lib().gsk_solver_freeze_limiters(self._h, C.c_int32(1 if on else 0)))

    def adjoint_prepare(self) -> None:
        """Tabulate the transposed Jacobian's tables at the current f (limiter branches recorded there)."""
        _check(lib().gsk_solver_adjoint_prepare(self._h))

    def adjoint_apply_device(self, lam_ptr: int, patch_weights, out_ptr: int, n: int) -> None:
        """out = J^T lam + (d/df sum_k w_k patch_mass_flow_k)^T (lam_ptr 0: lam = 0)."""
        w = (C.c_double * len(patch_weights))(*[float(x) for x in patch_weights])
        _check(lib().gsk_solver_adjoint_apply_device(self._h, C.c_void_p(lam_ptr or None), w, C.c_size_t(len(w)),
                                                     C.c_void_p(out_ptr), C.c_size_t(n)))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and _lib_handle is not None:
            _lib_handle.gsk_solver_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()
