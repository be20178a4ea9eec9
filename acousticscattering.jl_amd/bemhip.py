"""Host-side mirror of the reference's full-3D BEM API over the libbemhip C ABI.

Names, argument meaning and error behaviour follow brandynlucca/AcousticScattering.jl
(paths relative to that repository):

  bem(surface, boundary, k; incidence_angle=pi/2, ...)   src/solvers/bem/surface.jl:10-22
  solve_full_bem(boundary, k, quad; ...)                  src/solvers/bem/full.jl:52-66,273-287
  scattering_amplitude(sol; direction)                     src/solvers/bem/postprocessing.jl:1-13
  target_strength(f) = 20 log10 |f|                        src/postprocessing/target_strength.jl:8
  far_field(quad, xhat, k, p, dp)                          src/solvers/bem/full.jl:509-519
  frequency_sweep(solve, frequencies, c)                   src/postprocessing/sweeps.jl:103-110
  incidence_angle_sweep(surface, boundary, k, angles)      src/postprocessing/sweeps.jl:626-701
  mesh(nodes, triangles)                                   src/geometry/surface.jl:271-298
  Rigid / PressureRelease / FluidFilled(g, h)              src/core/materials.jl:16,23,62-72

Every solve runs on the GPU through libbemhip.so; there is no CPU path.
Importing this module fails loudly when the library has not been built.
"""
import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbemhip.so")

BEM_OK, BEM_ERR_ARGUMENT, BEM_ERR_DIMENSION, BEM_ERR_DEVICE = 0, 1, 2, 3
BEM_ERR_ADMISSION, BEM_ERR_NUMERIC, BEM_ERR_STATE = 4, 5, 6
BEM_RIGID, BEM_PRESSURE_RELEASE, BEM_FLUID = 0, 1, 2
BEM_BURTON_MILLER, BEM_CBIE, BEM_MULLER = 0, 1, 2
BEM_DOUBLE, BEM_MIXED = 0, 1
_PRECISION = {"double": BEM_DOUBLE, "mixed": BEM_MIXED}
REASONS = ["none", "conversion", "single_singular", "condition", "nonfinite",
           "iteration_limit", "stagnation"]

#: Every symbol declared in include/bemhip.h.
EXPORTED = [
    "bem_version", "bem_problem_defaults", "bem_ctx_create", "bem_ctx_destroy", "bem_last_error",
    "bem_ctx_memory", "bem_ctx_set_profiling", "bem_ctx_kernel_stats", "bem_ctx_reset_stats", "bem_ctx_set_lanes",
    "bem_ctx_set_heavy_lanes", "bem_ctx_set_lu_tuning", "bem_ctx_autotune", "bem_ctx_lu_tune",
    "bem_icosphere_size", "bem_icosphere", "bem_revolution_size", "bem_revolution",
    "bem_mesh_create", "bem_mesh_validate", "bem_mesh_destroy", "bem_mesh_info", "bem_mesh_panels",
    "bem_mesh_quadrature", "bem_mesh_near_count", "bem_operator_create", "bem_operator_destroy",
    "bem_operator_matrix", "bem_operator_size", "bem_operator_solve", "bem_solve",
    "bem_solution_destroy", "bem_solution_count", "bem_solution_traces", "bem_far_field",
    "bem_solution_diagnostics", "bem_sweep", "bem_server_create", "bem_server_destroy",
    "bem_server_submit", "bem_server_poll", "bem_server_wait", "bem_server_error",
    "bem_server_stats_get", "bem_hmatrix_defaults", "bem_hmatrix_plan", "bem_hmatrix_create",
    "bem_hmatrix_destroy", "bem_hmatrix_get_info", "bem_hmatrix_matvec", "bem_hmatrix_apply", "bem_hmatrix_solve",
    "bem_debug_lu", "bem_far_field_noalloc", "bem_pressure", "bem_reduced_defaults", "bem_reduced_sweep",
    "bem_axi_operators", "bem_axi_solve", "bem_axi_info", "bem_axi_traces", "bem_axi_far_field", "bem_axi_destroy",
    "bem_curved_icosphere_size", "bem_curved_icosphere", "bem_mesh_create_nystrom", "bem_nystrom_quadrature",
    "bem_mesh_nystrom_info",
]


class LuTune(ctypes.Structure):
    """bem_lu_tune: the shape of one factorization and the tuner's timings."""
    _fields_ = [("unknowns", ctypes.c_int64), ("precision", ctypes.c_int32), ("outer", ctypes.c_int32),
                ("group", ctypes.c_int32), ("split", ctypes.c_int32), ("ahead", ctypes.c_int32),
                ("graph", ctypes.c_int32), ("tuned", ctypes.c_int32), ("candidates", ctypes.c_int32),
                ("best_ms", ctypes.c_double), ("default_ms", ctypes.c_double)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["precision"] = "mixed" if self.precision == BEM_MIXED else "double"
        for k in ("split", "ahead", "graph", "tuned"):
            d[k] = bool(d[k])
        return d


class Problem(ctypes.Structure):
    _fields_ = [("boundary", ctypes.c_int32), ("formulation", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("equilibrate", ctypes.c_int32),
                ("k", ctypes.c_double), ("density_contrast", ctypes.c_double),
                ("soundspeed_contrast", ctypes.c_double), ("refine_tolerance", ctypes.c_double),
                ("refine_maxiter", ctypes.c_int32), ("force_promotion", ctypes.c_int32),
                ("condition_guard", ctypes.c_double)]


class Diagnostics(ctypes.Structure):
    _fields_ = [("unknown_count", ctypes.c_int64), ("quadrature_nodes", ctypes.c_int64),
                ("absolute_residual", ctypes.c_double), ("relative_residual", ctypes.c_double),
                ("scaled_absolute_residual", ctypes.c_double),
                ("scaled_relative_residual", ctypes.c_double),
                ("factor_precision", ctypes.c_int32), ("refinement_reason", ctypes.c_int32),
                ("refinement_iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("normwise_backward_error", ctypes.c_double),
                ("componentwise_backward_error", ctypes.c_double),
                ("single_reciprocal_condition", ctypes.c_double),
                ("assembly_ms", ctypes.c_double), ("factor_ms", ctypes.c_double),
                ("solve_ms", ctypes.c_double), ("iterations", ctypes.c_int32), ("method", ctypes.c_int32),
                ("derivative_evaluation", ctypes.c_int32)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["factor_precision"] = "double" if self.factor_precision == 0 else "single"
        d["method"] = ["direct", "mixed_refinement", "gmres"][self.method]
        # full.jl:336-339: how K' and H of each medium were evaluated (Mueller only)
        ev = {1: "direct", 2: "calderon"}.get(self.derivative_evaluation)
        d["derivative_evaluation"] = None if ev is None else {"exterior": ev, "interior": ev}
        d["refinement_reason"] = REASONS[self.refinement_reason]
        d["converged"] = bool(self.converged)
        return d


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_LIB = None


class ServerOptions(ctypes.Structure):
    _fields_ = [("lanes", ctypes.c_int32), ("keep_factors", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("queued", ctypes.c_int64), ("running", ctypes.c_int64),
                ("resident_operators", ctypes.c_int64), ("factor_reuses", ctypes.c_int64),
                ("resident_bytes", ctypes.c_uint64)]


class HMatrixOptions(ctypes.Structure):
    _fields_ = [("leaf_size", ctypes.c_int32), ("eta", ctypes.c_double), ("aca_tolerance", ctypes.c_double),
                ("max_rank", ctypes.c_int32), ("gmres_restart", ctypes.c_int32), ("gmres_maxiter", ctypes.c_int32),
                ("gmres_tolerance", ctypes.c_double), ("preconditioner", ctypes.c_int32),
                ("recycle_dimension", ctypes.c_int32)]


class HMatrixInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nodes", ctypes.c_int64), ("leaves", ctypes.c_int64),
                ("dense_blocks", ctypes.c_int64), ("lowrank_blocks", ctypes.c_int64), ("near_pairs", ctypes.c_int64),
                ("dense_bytes", ctypes.c_uint64), ("lowrank_bytes", ctypes.c_uint64), ("max_rank", ctypes.c_int32),
                ("unconverged_blocks", ctypes.c_int32), ("mean_rank", ctypes.c_double), ("plan_ms", ctypes.c_double),
                ("assembly_ms", ctypes.c_double), ("aca_ms", ctypes.c_double)]


class ReducedOptions(ctypes.Structure):
    _fields_ = [("tolerance", ctypes.c_double), ("rank_tolerance", ctypes.c_double), ("capacity", ctypes.c_int32),
                ("maxbytes", ctypes.c_int64)]


class ReducedReport(ctypes.Structure):
    _fields_ = [("k", ctypes.c_double), ("reduced", ctypes.c_int32), ("reason", ctypes.c_int32),
                ("rank", ctypes.c_int32), ("kept", ctypes.c_int32), ("projected_original", ctypes.c_double),
                ("projected_scaled", ctypes.c_double), ("residual_original", ctypes.c_double),
                ("residual_scaled", ctypes.c_double)]

    REASONS = {0: "accepted", 1: "empty", 2: "residual", 3: "rank", 4: "nonfinite"}

    def as_dict(self, training=False):
        d = {"k": self.k, "rank": self.rank, "residual": {"original": self.residual_original,
                                                          "scaled": self.residual_scaled}}
        if training:
            d["kept"] = bool(self.kept)
        else:
            d.update(reduced=bool(self.reduced), reason=self.REASONS[self.reason],
                     projected={"original": self.projected_original, "scaled": self.projected_scaled})
        return d


def load_library(path=LIB_PATH):
    """Load libbemhip.so (raises when it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise ImportError(f"libbemhip.so not built at {path}; run __graft_entry__.build()")
    L = ctypes.CDLL(path)
    sig = {
        "bem_version": (ctypes.c_char_p, []),
        "bem_problem_defaults": (None, [ctypes.POINTER(Problem), ctypes.c_int32]),
        "bem_ctx_create": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint64, ctypes.POINTER(_vp)]),
        "bem_ctx_destroy": (ctypes.c_int, [_vp]),
        "bem_last_error": (ctypes.c_char_p, [_vp]),
        "bem_ctx_memory": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
        "bem_ctx_set_profiling": (ctypes.c_int, [_vp, ctypes.c_int32]),
        "bem_ctx_kernel_stats": (ctypes.c_int, [_vp, ctypes.c_char_p, _ip, _dp, _dp, _dp]),
        "bem_ctx_reset_stats": (ctypes.c_int, [_vp]),
        "bem_ctx_set_lanes": (ctypes.c_int, [_vp, ctypes.c_int32]),
        "bem_ctx_set_heavy_lanes": (ctypes.c_int, [_vp, ctypes.c_int32]),
        "bem_ctx_set_lu_tuning": (ctypes.c_int, [_vp, ctypes.c_int32]),
        "bem_axi_operators": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp, ctypes.c_double, ctypes.c_int32,
                                             ctypes.c_int32, _dp, _dp]),
        "bem_axi_solve": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp, ctypes.c_void_p, ctypes.c_double,
                                         ctypes.c_int32, ctypes.POINTER(_vp)]),
        "bem_axi_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                        _dp, _dp, _dp]),
        "bem_axi_traces": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, _dp]),
        "bem_axi_far_field": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, _dp]),
        "bem_axi_destroy": (ctypes.c_int, [_vp]),
        "bem_ctx_autotune": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, _vp]),
        "bem_ctx_lu_tune": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_int32, _vp]),
        "bem_icosphere_size": (ctypes.c_int, [ctypes.c_int32, _ip, _ip]),
        "bem_icosphere": (ctypes.c_int, [ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _dp, _ip]),
        "bem_revolution_size": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _ip, _ip]),
        "bem_revolution": (ctypes.c_int, [ctypes.c_int32, _dp, _dp, ctypes.c_int32, _dp, _ip]),
        "bem_mesh_create": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, ctypes.c_int64, _ip, ctypes.POINTER(_vp)]),
        "bem_curved_icosphere_size": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _ip, _ip]),
        "bem_curved_icosphere": (ctypes.c_int, [ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int32, _dp, _ip]),
        "bem_mesh_create_nystrom": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int32, _ip,
                                                   ctypes.c_int32, ctypes.POINTER(_vp)]),
        "bem_nystrom_quadrature": (ctypes.c_int, [ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int32, _ip,
                                                  ctypes.c_int32, _ip, _dp, _dp, _dp, ctypes.c_char_p,
                                                  ctypes.c_int32]),
        "bem_mesh_nystrom_info": (ctypes.c_int, [_vp, _ip, ctypes.POINTER(ctypes.c_int32),
                                                 ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
        "bem_mesh_validate": (ctypes.c_int, [ctypes.c_int64, _dp, ctypes.c_int64, _ip, _dp,
                                             ctypes.c_char_p, ctypes.c_int32]),
        "bem_mesh_destroy": (ctypes.c_int, [_vp]),
        "bem_mesh_info": (ctypes.c_int, [_vp, _ip, _ip, _dp, _dp]),
        "bem_mesh_panels": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
        "bem_mesh_quadrature": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
        "bem_mesh_near_count": (ctypes.c_int, [_vp, _ip]),
        "bem_operator_create": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(Problem), ctypes.POINTER(_vp)]),
        "bem_operator_destroy": (ctypes.c_int, [_vp]),
        "bem_operator_matrix": (ctypes.c_int, [_vp, _dp, ctypes.c_int64]),
        "bem_operator_size": (ctypes.c_int, [_vp, _ip]),
        "bem_operator_solve": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, ctypes.POINTER(_vp)]),
        "bem_solve": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(Problem), ctypes.c_double,
                                     ctypes.c_double, ctypes.POINTER(_vp)]),
        "bem_solution_destroy": (ctypes.c_int, [_vp]),
        "bem_solution_count": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int32), _ip]),
        "bem_solution_traces": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, _dp]),
        "bem_far_field": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int64, _dp, _dp]),
        "bem_far_field_noalloc": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int64, _dp, _dp]),
        "bem_pressure": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int64, _dp, ctypes.c_int32, _dp]),
        "bem_reduced_defaults": (None, [ctypes.POINTER(ReducedOptions)]),
        "bem_reduced_sweep": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_int32, _dp,
                                             ctypes.c_int32, _dp, ctypes.c_double, ctypes.c_double,
                                             ctypes.POINTER(ReducedOptions), _dp, ctypes.POINTER(ReducedReport),
                                             ctypes.POINTER(ReducedReport), ctypes.POINTER(ctypes.c_int32)]),
        "bem_solution_diagnostics": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(Diagnostics)]),
        "bem_sweep": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.POINTER(_vp), ctypes.POINTER(Problem),
                                     _dp, _dp, _dp, ctypes.POINTER(Diagnostics)]),
        "bem_debug_lu": (ctypes.c_int, [_vp, ctypes.c_int64, _dp, ctypes.c_int32, _dp, _ip, _dp, _dp]),
        "bem_server_create": (ctypes.c_int, [_vp, ctypes.POINTER(ServerOptions), ctypes.POINTER(_vp)]),
        "bem_server_destroy": (ctypes.c_int, [_vp]),
        "bem_server_submit": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(Problem), ctypes.c_double,
                                             ctypes.c_double, _ip]),
        "bem_server_poll": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32)]),
        "bem_server_wait": (ctypes.c_int, [_vp, ctypes.c_int64, ctypes.c_double, _dp,
                                           ctypes.POINTER(Diagnostics), _dp]),
        "bem_server_error": (ctypes.c_char_p, [_vp, ctypes.c_int64]),
        "bem_server_stats_get": (ctypes.c_int, [_vp, ctypes.POINTER(ServerStats)]),
        "bem_hmatrix_defaults": (None, [ctypes.POINTER(HMatrixOptions)]),
        "bem_hmatrix_plan": (ctypes.c_int, [ctypes.c_int64, _dp, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                            _ip, _ip, _ip, _ip, _ip, ctypes.c_int64]),
        "bem_hmatrix_create": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(Problem), ctypes.POINTER(HMatrixOptions),
                                              ctypes.POINTER(_vp)]),
        "bem_hmatrix_destroy": (ctypes.c_int, [_vp]),
        "bem_hmatrix_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(HMatrixInfo)]),
        "bem_hmatrix_matvec": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, _dp]),
        "bem_hmatrix_apply": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp]),
        "bem_hmatrix_solve": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, ctypes.POINTER(_vp)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class DimensionMismatch(ValueError):
    pass


class AdmissionError(RuntimeError):
    pass


def _raise(code, msg):
    if code == BEM_ERR_ARGUMENT:
        raise ValueError(msg)          # ArgumentError
    if code == BEM_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if code == BEM_ERR_ADMISSION:
        raise AdmissionError(msg)
    raise RuntimeError(msg)            # ErrorException


# ---------------------------------------------------------------------------
# boundary conditions and bodies
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class Rigid:
    code: int = BEM_RIGID


@dataclass(frozen=True)
class PressureRelease:
    code: int = BEM_PRESSURE_RELEASE


@dataclass(frozen=True)
class FluidFilled:
    """g = rho_int / rho_ext, h = c_int / c_ext (src/core/materials.jl:62-72)."""
    density_contrast: float
    soundspeed_contrast: float
    code: int = BEM_FLUID


@dataclass(frozen=True)
class Sphere:
    radius: float


@dataclass(frozen=True)
class Spheroid:
    """Semi-axis `a` along x, `b` along y and z (src/geometry/spheroid.jl:7-28)."""
    a: float
    b: float


# ---------------------------------------------------------------------------
# context
# ---------------------------------------------------------------------------
class Context:
    """One GPU: a HIP stream, an HBM budget and the library's resident state."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        L = load_library()
        h = _vp()
        rc = L.bem_ctx_create(device, hbm_budget_bytes, ctypes.byref(h))
        if rc != BEM_OK:
            _raise(rc, f"bem_ctx_create failed on device {device} (code {rc})")
        self._h = h
        self.lib = L

    def check(self, rc):
        if rc != BEM_OK:
            _raise(rc, self.lib.bem_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self.lib.bem_ctx_destroy(self._h)
            self._h = None

    def memory(self):
        used, budget = ctypes.c_uint64(), ctypes.c_uint64()
        self.check(self.lib.bem_ctx_memory(self._h, ctypes.byref(used), ctypes.byref(budget)))
        return used.value, budget.value

    def set_profiling(self, on=True):
        self.check(self.lib.bem_ctx_set_profiling(self._h, 1 if on else 0))

    def reset_stats(self):
        self.check(self.lib.bem_ctx_reset_stats(self._h))

    def set_lanes(self, max_lanes):
        """Largest number of concurrent solves of batched sweeps (1..8)."""
        self.check(self.lib.bem_ctx_set_lanes(self._h, int(max_lanes)))

    def set_heavy_lanes(self, max_heavy):
        """Largest number of concurrent sweep solves with 16384 or more
        unknowns (1..8); the other lanes take smaller solves meanwhile."""
        self.check(self.lib.bem_ctx_set_heavy_lanes(self._h, int(max_heavy)))

    def set_lu_tuning(self, mode):
        """"default" (outer 512, FP32 3M updates, look-ahead), "split" (the
        default with bf16-split outer updates for complex FP32) or "tuned"
        (per size class the tuner's choice; the default)."""
        m = {"default": 0, "split": 1, "tuned": 2}
        if mode not in m:
            raise ValueError("mode must be 'default', 'split' or 'tuned'")
        self.check(self.lib.bem_ctx_set_lu_tuning(self._h, m[mode]))

    def autotune(self, n, precision="mixed"):
        """Time candidate factorization shapes for n unknowns and keep the
        fastest for that size class; returns the tuner's report."""
        r = LuTune()
        self.check(self.lib.bem_ctx_autotune(self._h, int(n), _PRECISION[precision], ctypes.byref(r)))
        return r.as_dict()

    def lu_tune(self, n, precision="mixed"):
        """The factorization shape n unknowns would use now."""
        r = LuTune()
        self.check(self.lib.bem_ctx_lu_tune(self._h, int(n), _PRECISION[precision], ctypes.byref(r)))
        return r.as_dict()

    def kernel_stats(self, name):
        n = ctypes.c_int64()
        ms, fl, by = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self.check(self.lib.bem_ctx_kernel_stats(self._h, name.encode(), ctypes.byref(n),
                                                 ctypes.byref(ms), ctypes.byref(fl), ctypes.byref(by)))
        return {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}


_DEFAULT = None


def default_context():
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


# ---------------------------------------------------------------------------
# meshes
# ---------------------------------------------------------------------------
def curved_icosphere_arrays(level, ax=1.0, ay=None, az=None, order=2):
    """Level-`level` icosphere with order-`order` Lagrange triangles (extra nodes
    on the sphere), mapped by (ax, ay, az): nodes 3xNv, triangles npt x N
    (1-based, Gmsh order)."""
    L = load_library()
    ay = ax if ay is None else ay
    az = ay if az is None else az
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    if L.bem_curved_icosphere_size(level, order, ctypes.byref(nv), ctypes.byref(nt)) != BEM_OK:
        raise ValueError("mesh_order must be 1, 2 or 3")
    npt = {1: 3, 2: 6, 3: 10}[order]
    nodes = np.empty((nv.value, 3))
    tri = np.empty((nt.value, npt), dtype=np.int64)
    if L.bem_curved_icosphere(level, ax, ay, az, order, _d(nodes), _i(tri)) != BEM_OK:
        raise ValueError("invalid icosphere")
    return nodes.T.copy(), tri.T.copy()


def nystrom_quadrature(nodes, triangles, qorder=4):
    """Quadrature nodes of a (curved) triangle mesh on the host: (coords n x 3,
    normals n x 3, weights n) in the node order of the device unknowns
    (element by element).  Raises ValueError with the reference's messages."""
    L = load_library()
    nodes = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).T)
    tri = np.ascontiguousarray(np.asarray(triangles, dtype=np.int64).T)
    n = ctypes.c_int64()
    buf = ctypes.create_string_buffer(256)
    args = (nodes.shape[0], _d(nodes), tri.shape[0], tri.shape[1], _i(tri), int(qorder))
    if L.bem_nystrom_quadrature(*args, ctypes.byref(n), None, None, None, buf, 256) != BEM_OK:
        raise ValueError(buf.value.decode())
    xyz = np.empty((n.value, 3)); nr = np.empty((n.value, 3)); w = np.empty(n.value)
    if L.bem_nystrom_quadrature(*args, ctypes.byref(n), _d(xyz), _d(nr), _d(w), buf, 256) != BEM_OK:
        raise ValueError(buf.value.decode())
    return xyz, nr, w


def icosphere_arrays(level, ax=1.0, ay=None, az=None):
    """Level-`level` icosphere mapped by (ax, ay, az): nodes 3xNv, triangles 3xN (1-based)."""
    L = load_library()
    ay = ax if ay is None else ay
    az = ay if az is None else az
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    if L.bem_icosphere_size(level, ctypes.byref(nv), ctypes.byref(nt)) != BEM_OK:
        raise ValueError("invalid icosphere level")
    nodes = np.empty((nv.value, 3))
    tri = np.empty((nt.value, 3), dtype=np.int64)
    if L.bem_icosphere(level, ax, ay, az, _d(nodes), _i(tri)) != BEM_OK:
        raise ValueError("invalid icosphere axes")
    return nodes.T.copy(), tri.T.copy()


def revolution_arrays(xs, rs, n_phi):
    """Body of revolution about x from meridian samples (poles at both ends)."""
    L = load_library()
    xs = np.ascontiguousarray(xs, dtype=np.float64)
    rs = np.ascontiguousarray(rs, dtype=np.float64)
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    if L.bem_revolution_size(len(xs), n_phi, ctypes.byref(nv), ctypes.byref(nt)) != BEM_OK:
        raise ValueError("need at least 3 meridian samples and 3 azimuthal segments")
    nodes = np.empty((nv.value, 3))
    tri = np.empty((nt.value, 3), dtype=np.int64)
    if L.bem_revolution(len(xs), _d(xs), _d(rs), n_phi, _d(nodes), _i(tri)) != BEM_OK:
        raise ValueError("invalid meridian")
    return nodes.T.copy(), tri.T.copy()


def validate_surface(nodes, triangles):
    """Topology checks of src/geometry/surface.jl:341-405 (host only); returns resolution."""
    L = load_library()
    nodes = np.ascontiguousarray(np.asarray(nodes, dtype=np.float64).T)
    tri = np.ascontiguousarray(np.asarray(triangles, dtype=np.int64).T)
    res = ctypes.c_double()
    buf = ctypes.create_string_buffer(256)
    rc = L.bem_mesh_validate(nodes.shape[0], _d(nodes), tri.shape[0], _i(tri), ctypes.byref(res), buf, 256)
    if rc != BEM_OK:
        raise ValueError(buf.value.decode())
    return res.value


class Mesh:
    """A closed triangulated surface resident on the device (`Mesh` of src/geometry/mesh.jl:39-44).

    Flat 3-node triangles with no `qorder` use one unknown per panel
    (centroid collocation).  6- or 10-node triangles, or any triangles with a
    `qorder`, use the reference's Nystrom discretisation: an order-`qorder`
    quadrature on every curved element (default 4), whose nodes are the
    unknowns (src/geometry/surface.jl:271-298,331)."""

    def __init__(self, nodes, triangles, ctx=None, body=None, qorder=None):
        self.ctx = ctx or default_context()
        nodes = np.asarray(nodes, dtype=np.float64)
        triangles = np.asarray(triangles)
        if nodes.ndim != 2 or nodes.shape[0] != 3:
            raise ValueError("nodes must have three rows")
        if triangles.ndim != 2 or triangles.shape[0] not in (3, 6, 10):
            raise ValueError("triangles must have 3, 6 or 10 rows in Gmsh ordering")
        if not np.all(np.isfinite(nodes)):
            raise ValueError("node coordinates must be finite")
        self.nodes = nodes
        self.triangles = triangles.astype(np.int64)
        self.body = body
        self.nystrom = qorder is not None or triangles.shape[0] != 3
        self.qorder = (4 if qorder is None else int(qorder)) if self.nystrom else None
        h = _vp()
        nd = np.ascontiguousarray(nodes.T)
        tr = np.ascontiguousarray(self.triangles.T)
        if self.nystrom:
            if not 1 <= self.qorder <= 8:
                raise ValueError("qorder must be between 1 and 8")
            self.ctx.check(self.ctx.lib.bem_mesh_create_nystrom(self.ctx._h, nd.shape[0], _d(nd), tr.shape[0],
                                                                tr.shape[1], _i(tr), self.qorder, ctypes.byref(h)))
        else:
            self.ctx.check(self.ctx.lib.bem_mesh_create(self.ctx._h, nd.shape[0], _d(nd), tr.shape[0], _i(tr),
                                                        ctypes.byref(h)))
        self._h = h
        self.n_elements = tr.shape[0]
        nt, nv = ctypes.c_int64(), ctypes.c_int64()
        res, vol = ctypes.c_double(), ctypes.c_double()
        self.ctx.lib.bem_mesh_info(h, ctypes.byref(nt), ctypes.byref(nv), ctypes.byref(res), ctypes.byref(vol))
        self.n, self.resolution, self.volume = nt.value, res.value, vol.value

    def __len__(self):
        return self.n

    def __del__(self):
        if getattr(self, "_h", None) and self.ctx._h:
            self.ctx.lib.bem_mesh_destroy(self._h)
            self._h = None

    def panels(self):
        c = np.empty((self.n, 3)); nr = np.empty((self.n, 3)); a = np.empty(self.n)
        self.ctx.check(self.ctx.lib.bem_mesh_panels(self._h, _d(c), _d(nr), _d(a)))
        return c, nr, a

    def quadrature(self):
        """(coords, normals, weights), the shape far_field consumes: the 7-point
        Gauss nodes of every panel, or the Nystrom nodes (the unknowns)."""
        m = self.n if self.nystrom else 7 * self.n
        xyz = np.empty((m, 3)); nr = np.empty((m, 3)); w = np.empty(m)
        self.ctx.check(self.ctx.lib.bem_mesh_quadrature(self._h, _d(xyz), _d(nr), _d(w)))
        return xyz, nr, w

    def near_count(self):
        c = ctypes.c_int64()
        self.ctx.check(self.ctx.lib.bem_mesh_near_count(self._h, ctypes.byref(c)))
        return c.value


def mesh(nodes_or_body, triangles=None, *, level=3, n_phi=64, units="m", ctx=None, mesh_order=None,
         qorder=None):
    """`mesh(nodes, triangles; qorder)` for a supplied surface (1-based
    triangles with 3, 6 or 10 rows, src/geometry/surface.jl:271-298),
    `mesh(body; level, mesh_order, qorder)` for a Sphere / Spheroid meshed as
    a mapped icosphere (flat panels, or curved order-`mesh_order` Nystrom
    elements as gmsh_sphere_mesh / gmsh_spheroid_mesh build,
    src/geometry/canonical_mesh.jl:3-36), or `mesh(cylinder; n_phi)` for a
    (bent) capped Cylinder with n_phi azimuthal segments."""
    scale = {"m": 1.0, "cm": 1e-2, "mm": 1e-3}.get(units)
    if scale is None:
        raise ValueError("units must be :m, :cm or :mm")
    if triangles is not None:
        return Mesh(np.asarray(nodes_or_body, dtype=np.float64) * scale, triangles, ctx=ctx, qorder=qorder)
    body = nodes_or_body
    if mesh_order is not None or qorder is not None:
        order = 2 if mesh_order is None else mesh_order
        if order not in (1, 2, 3):
            raise ValueError("mesh_order must be 1, 2 or 3")
        if isinstance(body, Sphere):
            nodes, tri = curved_icosphere_arrays(level, body.radius * scale, order=order)
        elif isinstance(body, Spheroid):
            nodes, tri = curved_icosphere_arrays(level, body.a * scale, body.b * scale, body.b * scale, order=order)
        else:
            raise ValueError("curved meshes are built for spheres and spheroids")
        return Mesh(nodes, tri, ctx=ctx, body=body, qorder=4 if qorder is None else qorder)
    if isinstance(body, Sphere):
        nodes, tri = icosphere_arrays(level, body.radius * scale)
    elif isinstance(body, Spheroid):
        nodes, tri = icosphere_arrays(level, body.a * scale, body.b * scale, body.b * scale)
    elif isinstance(body, Cylinder):
        nodes, tri = cylinder_arrays(Cylinder(body.radius * scale, body.length * scale,
                                              body.radius_curvature * scale, body.endcap_depth * scale), n_phi)
    else:
        raise ValueError("unsupported body")
    return Mesh(nodes, tri, ctx=ctx, body=body)


# ---------------------------------------------------------------------------
# solves
# ---------------------------------------------------------------------------
def incidence_direction(beta, alpha=0.0):
    """Polar angle from +x; azimuth from +y toward +z (src/solvers/bem/full.jl:15-18)."""
    return np.array([math.cos(beta), math.sin(beta) * math.cos(alpha), math.sin(beta) * math.sin(alpha)])


_FORMULATIONS = {"burton_miller": BEM_BURTON_MILLER, "cbie": BEM_CBIE, "muller": BEM_MULLER}


def make_problem(boundary, k, *, formulation=None, precision="double", refinement=None,
                 equilibrate=True, _force_promotion=False):
    L = load_library()
    p = Problem()
    L.bem_problem_defaults(ctypes.byref(p), boundary.code)
    p.k = float(k)
    if formulation is not None:
        if formulation not in _FORMULATIONS:
            raise ValueError("formulation must be :cbie or :burton_miller" if boundary.code != BEM_FLUID
                             else "fluid formulation must be :muller or :cbie")
        p.formulation = _FORMULATIONS[formulation]
    if isinstance(boundary, FluidFilled):
        p.density_contrast = float(boundary.density_contrast)
        p.soundspeed_contrast = float(boundary.soundspeed_contrast)
    if precision not in ("double", "mixed"):
        raise ValueError("precision must be :double or :mixed")
    p.precision = BEM_MIXED if precision == "mixed" else BEM_DOUBLE
    p.equilibrate = 1 if equilibrate else 0
    refinement = dict(refinement or {})
    if refinement and precision != "mixed":
        raise ValueError("refinement options require precision=:mixed")
    for key in refinement:
        if key not in ("tolerance", "maxiter", "condition_guard"):
            raise ValueError("unknown refinement option")
    if "tolerance" in refinement:
        p.refine_tolerance = float(refinement["tolerance"])
    if "maxiter" in refinement:
        mi = refinement["maxiter"]
        if not isinstance(mi, int) or mi < 0:
            raise ValueError("maxiter must be a nonnegative integer")
        p.refine_maxiter = mi
    if "condition_guard" in refinement:
        p.condition_guard = float(refinement["condition_guard"])
    p.force_promotion = 1 if _force_promotion else 0
    return p


class Operator:
    """Assembled and factored system resident in HBM; solve() for many incidences."""

    def __init__(self, surface, boundary, k, **kw):
        self.mesh = surface
        self.ctx = surface.ctx
        self.boundary = boundary
        self.k = float(k)
        self.problem = make_problem(boundary, k, **kw)
        h = _vp()
        self.ctx.check(self.ctx.lib.bem_operator_create(self.ctx._h, surface._h, ctypes.byref(self.problem),
                                                        ctypes.byref(h)))
        self._h = h
        n = ctypes.c_int64()
        self.ctx.lib.bem_operator_size(h, ctypes.byref(n))
        self.n = n.value

    def __del__(self):
        if getattr(self, "_h", None) and self.ctx._h:
            self.ctx.lib.bem_operator_destroy(self._h)
            self._h = None

    def matrix(self):
        A = np.empty((self.n, self.n), dtype=np.complex128, order="F")
        self.ctx.check(self.ctx.lib.bem_operator_matrix(self._h, A.ctypes.data_as(_dp), self.n))
        return A

    def solve(self, angles):
        """angles: sequence of (incidence_angle, incidence_azimuth)."""
        ang = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).reshape(-1, 2))
        h = _vp()
        self.ctx.check(self.ctx.lib.bem_operator_solve(self._h, ang.shape[0], _d(ang), ctypes.byref(h)))
        return _Solution(self.ctx, h, self.mesh, self.k, ang)


class _Solution:
    def __init__(self, ctx, h, mesh, k, angles):
        self.ctx, self._h, self.mesh, self.k, self.angles = ctx, h, mesh, k, angles

    def __del__(self):
        if getattr(self, "_h", None) and self.ctx._h:
            self.ctx.lib.bem_solution_destroy(self._h)
            self._h = None

    def traces(self, r=0):
        p = np.empty(self.mesh.n, dtype=np.complex128)
        q = np.empty(self.mesh.n, dtype=np.complex128)
        self.ctx.check(self.ctx.lib.bem_solution_traces(self._h, r, p.ctypes.data_as(_dp), q.ctypes.data_as(_dp)))
        return p, q

    def far_field(self, r=0, directions=None):
        if directions is None:
            out = np.empty(1, dtype=np.complex128)
            self.ctx.check(self.ctx.lib.bem_far_field(self._h, r, 1, None, out.ctypes.data_as(_dp)))
            return out
        dirs = np.ascontiguousarray(np.atleast_2d(np.asarray(directions, dtype=np.float64)))
        if dirs.shape[1] != 3:
            raise ValueError("direction must be a finite unit vector with three components")
        out = np.empty(dirs.shape[0], dtype=np.complex128)
        self.ctx.check(self.ctx.lib.bem_far_field(self._h, r, dirs.shape[0], _d(dirs), out.ctypes.data_as(_dp)))
        return out

    def pressure(self, points, field="total", r=0):
        """Complex pressure at points (3 x n, or a sequence of 3-vectors)."""
        codes = {"total": 0, "scattered": 1, "incident": 2, "interior": 3}
        if field not in codes:
            raise ValueError("field must be :total, :scattered, :incident or :interior")
        pts = np.asarray(points, dtype=np.float64)
        if pts.ndim == 1:
            pts = pts.reshape(3, 1)
        elif pts.shape[0] != 3 and pts.shape[-1] == 3:
            pts = pts.T
        if pts.shape[0] != 3:
            raise ValueError("points must have three coordinates")
        P = np.ascontiguousarray(pts.T)
        out = np.empty(P.shape[0], dtype=np.complex128)
        self.ctx.check(self.ctx.lib.bem_pressure(self._h, r, P.shape[0], _d(P), codes[field],
                                                 out.ctypes.data_as(_dp)))
        return out

    def diagnostics(self, r=0):
        d = Diagnostics()
        self.ctx.check(self.ctx.lib.bem_solution_diagnostics(self._h, r, ctypes.byref(d)))
        return d.as_dict()


def solve_full_bem(boundary, k, surface, *, incidence_angle=0.0, incidence_azimuth=0.0, incident=None,
                   formulation=None, compression=None, correction=None, gmres_kwargs=None,
                   equilibrate=True, condition_limit=512, precision="double", refinement=None,
                   return_diagnostics=False):
    """Drop-in for solve_full_bem (src/solvers/bem/full.jl:52-66,273-287).

    Returns (p_scat, dpdn_scat, quad[, diagnostics]) at the panel collocation
    points.  Without compression the dense direct solve replaces the
    reference's GMRES for rigid/soft (its converged limit).  With
    compression={"method": "hmatrix", ...} a rigid surface or a fluid body
    (Mueller) is solved by the compressed operator and GMRES (HMatrix; the
    remaining compression keys are HMatrix options, gmres_kwargs may set
    reltol / restart / maxiter).
    `correction` is accepted for signature compatibility only.
    """
    if incident is not None:
        raise ValueError("only the plain unit plane wave is supported on the device path")
    if not (math.isfinite(incidence_angle) and math.isfinite(incidence_azimuth)):
        raise ValueError("incidence angles must be finite")
    if condition_limit < 0:
        raise ValueError("condition_limit must be nonnegative")
    method = (compression or {}).get("method", "none")
    if method not in ("none", "hmatrix"):
        raise ValueError("compression method must be :none or :hmatrix")
    if method == "hmatrix":
        opts = {key: v for key, v in compression.items() if key != "method"}
        g = dict(gmres_kwargs or {})
        if "reltol" in g:
            opts["gmres_tolerance"] = g.pop("reltol")
        if "restart" in g:
            opts["gmres_restart"] = g.pop("restart")
        if "maxiter" in g:
            opts["gmres_maxiter"] = g.pop("maxiter")
        H = HMatrix(surface, boundary, k, formulation=formulation, **opts)
        sol = H.solve([(incidence_angle, incidence_azimuth)])
        p, q = sol.traces(0)
        if return_diagnostics:
            d = sol.diagnostics(0)
            d["formulation"] = {BEM_BURTON_MILLER: "burton_miller", BEM_CBIE: "cbie",
                                BEM_MULLER: "muller"}[H.problem.formulation]
            d["compression"] = dict(compression)
            return p, q, surface, d
        return p, q, surface
    op = Operator(surface, boundary, k, formulation=formulation, precision=precision,
                  refinement=refinement, equilibrate=equilibrate)
    sol = op.solve([(incidence_angle, incidence_azimuth)])
    p, q = sol.traces(0)
    if return_diagnostics:
        d = sol.diagnostics(0)
        d["formulation"] = {BEM_BURTON_MILLER: "burton_miller", BEM_CBIE: "cbie",
                            BEM_MULLER: "muller"}[op.problem.formulation]
        d["method"] = "mixed_refinement" if precision == "mixed" else "direct"
        return p, q, surface, d
    return p, q, surface


@dataclass
class BEMSolution:
    """Result of bem(); fields follow src/solvers/bem/types.jl:2-36."""
    body: object
    boundary: object
    k: float
    method: str
    surface: Mesh
    p_scat: np.ndarray
    dpdn_scat: np.ndarray
    incidence_angle: float
    incidence_azimuth: float
    diagnostics: dict = field(default_factory=dict)
    _solution: object = None


# ---------------------------------------------------------------------------
# axisymmetric path (src/solvers/bem/axisymmetric.jl)
# ---------------------------------------------------------------------------
class MeridianMesh:
    """MeridianMesh(rho, z) (src/geometry/meridian.jl:16-19): the generating
    curve of a body of revolution, north pole to south pole, z along the axis
    of symmetry (Cartesian x of the 3D path)."""

    def __init__(self, rho, z):
        self.rho = np.ascontiguousarray(np.asarray(rho, dtype=np.float64))
        self.z = np.ascontiguousarray(np.asarray(z, dtype=np.float64))
        if self.rho.shape != self.z.shape or self.rho.ndim != 1:
            raise ValueError("a meridian needs matching rho/z vectors")

    @property
    def n(self):
        return len(self.rho) - 1


def sphere_mesh(a, n):
    """src/geometry/meridian.jl:201-205: n panels at uniform polar angle."""
    if n < 3:
        raise ValueError("n must be at least 3 (one panel per segment)")
    psi = np.linspace(0.0, math.pi, n + 1)
    return MeridianMesh(a * np.sin(psi), a * np.cos(psi))


def spheroid_mesh(a, b, n):
    """src/geometry/meridian.jl:238-259: semi-axis a along the axis, b
    equatorial, nodes at equal arc length along the meridian ellipse."""
    if n < 3:
        raise ValueError("n must be at least 3 (one panel per segment)")
    psi = np.linspace(0.0, math.pi, max(2000, 40 * n))
    speed = np.hypot(b * np.cos(psi), a * np.sin(psi))
    arc = np.concatenate([[0.0], np.cumsum(0.5 * (speed[1:] + speed[:-1]) * np.diff(psi))])
    nodes = np.interp(np.linspace(0.0, arc[-1], n + 1), arc, psi)
    nodes[-1] = math.pi
    return MeridianMesh(b * np.sin(nodes), a * np.cos(nodes))


def _characteristic_radius(body):
    if isinstance(body, Sphere):
        return body.radius
    if isinstance(body, Spheroid):
        return max(body.a, body.b)
    raise ValueError("the axisymmetric path takes a Sphere or a Spheroid")


def bem_panel_count(k, characteristic_radius, elements_per_wavelength=30):
    """src/geometry/meridian.jl:190-192."""
    return max(20, math.ceil(elements_per_wavelength * k * characteristic_radius / 2))


def default_mode_count(ka):
    """src/solvers/modal/sphere.jl:1-3."""
    return max(20, math.ceil(2.5 * ka + 4.0 * np.cbrt(ka + 1.0) + 10.0))


def axisymmetric_mesh(body, n):
    """_axisymmetric_mesh (src/geometry/mesh.jl:24-26)."""
    if isinstance(body, Sphere):
        return sphere_mesh(body.radius, n)
    if isinstance(body, Spheroid):
        return spheroid_mesh(body.a, body.b, n)
    raise ValueError("the axisymmetric path takes a Sphere or a Spheroid")


class AxisymmetricSolution:
    """BEMSolution{_AxisymmetricSurfaceData}: per-mode scattered traces on a
    meridian, far field by the mode sum (src/postprocessing/farfield.jl)."""

    def __init__(self, ctx, h, body, boundary, k, mesh, incidence_angle, m_max):
        self.ctx, self._h = ctx, h
        self.body, self.boundary, self.k = body, boundary, float(k)
        self.method = "axisymmetric"
        self.mesh = mesh
        self.incidence_angle = float(incidence_angle)
        npan, nm = ctypes.c_int32(), ctypes.c_int32()
        res, ta, ts = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        ctx.check(ctx.lib.bem_axi_info(h, ctypes.byref(npan), ctypes.byref(nm), ctypes.byref(res), ctypes.byref(ta),
                                       ctypes.byref(ts)))
        self.n, self.n_modes = npan.value, nm.value
        self.diagnostics = {"method": "axisymmetric", "relative_residual": res.value, "converged": None,
                            "iterations": None, "modes": list(range(self.n_modes)),
                            "solver_options": {"n": self.n, "incidence_angle": self.incidence_angle,
                                               "m_max": m_max},
                            "assembly_ms": ta.value, "solve_ms": ts.value}

    def __del__(self):
        if getattr(self, "_h", None) and self.ctx._h:
            self.ctx.lib.bem_axi_destroy(self._h)
            self._h = None

    def traces(self, mode=0):
        """(p_scat, dpdn_scat) of Fourier mode `mode` on the panels."""
        p = np.empty(self.n, dtype=np.complex128)
        d = np.empty(self.n, dtype=np.complex128)
        self.ctx.check(self.ctx.lib.bem_axi_traces(self._h, int(mode), p.ctypes.data_as(_dp), d.ctypes.data_as(_dp)))
        return p, d

    def far_field(self, angles):
        a = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).reshape(-1, 2))
        out = np.empty(a.shape[0], dtype=np.complex128)
        self.ctx.check(self.ctx.lib.bem_axi_far_field(self._h, a.shape[0], a.ctypes.data_as(_dp),
                                                      out.ctypes.data_as(_dp)))
        return out


def _bem_axisymmetric(body_or_mesh, boundary, k, incidence_angle, n=None, m_max=None, ctx=None):
    if not isinstance(boundary, (Rigid, PressureRelease, FluidFilled)):
        raise ValueError("the axisymmetric path solves rigid, pressure-release and fluid-filled bodies")
    if not math.isfinite(incidence_angle):
        raise ValueError("k must be positive and finite and incidence angles finite")
    if isinstance(body_or_mesh, MeridianMesh):
        mesh, body = body_or_mesh, None
        radius = float(np.max(np.hypot(mesh.rho, mesh.z - 0.5 * (mesh.z.max() + mesh.z.min()))))
    else:
        body = body_or_mesh
        radius = _characteristic_radius(body)
        mesh = axisymmetric_mesh(body, int(n) if n is not None else bem_panel_count(k, radius))
    if m_max is None:
        m_max = default_mode_count(k * radius)
    if m_max < 0:
        raise ValueError("m_max must be nonnegative")
    ctx = ctx or default_context()
    prob = make_problem(boundary, k)
    h = _vp()
    ctx.check(ctx.lib.bem_axi_solve(ctx._h, len(mesh.rho), mesh.rho.ctypes.data_as(_dp), mesh.z.ctypes.data_as(_dp),
                                    ctypes.byref(prob), float(incidence_angle), int(m_max), ctypes.byref(h)))
    return AxisymmetricSolution(ctx, h, body, boundary, k, mesh, incidence_angle, int(m_max))


def axisymmetric_operators(mesh, k, modes, ctx=None):
    """assemble_cbie_operators(mesh, k; m) for m in `modes` (a contiguous
    range): returns (K, V), arrays of shape (len(modes), n, n)."""
    modes = list(modes)
    if not modes or modes != list(range(modes[0], modes[0] + len(modes))):
        raise ValueError("modes must be a nonempty contiguous range")
    ctx = ctx or default_context()
    n = mesh.n
    K = np.empty((len(modes), n, n), dtype=np.complex128, order="C")
    V = np.empty((len(modes), n, n), dtype=np.complex128, order="C")
    Kf = np.empty(len(modes) * n * n, dtype=np.complex128)
    Vf = np.empty(len(modes) * n * n, dtype=np.complex128)
    ctx.check(ctx.lib.bem_axi_operators(ctx._h, len(mesh.rho), mesh.rho.ctypes.data_as(_dp),
                                        mesh.z.ctypes.data_as(_dp), float(k), modes[0], len(modes),
                                        Kf.ctypes.data_as(_dp), Vf.ctypes.data_as(_dp)))
    for c in range(len(modes)):
        K[c] = Kf[c * n * n:(c + 1) * n * n].reshape(n, n, order="F")
        V[c] = Vf[c * n * n:(c + 1) * n * n].reshape(n, n, order="F")
    return K, V


def bem(surface, boundary, k, *, incidence_angle=math.pi / 2, incidence_azimuth=0.0, **kw):
    """Drop-in for bem(surface::Mesh, boundary, k; ...) (src/solvers/bem/surface.jl:10-22)
    and bem(body, boundary, k; method=:axisymmetric, n, m_max) for a Sphere or
    Spheroid (src/solvers/bem/dispatch.jl:79-104; axisymmetric by default, as
    there) or a MeridianMesh."""
    if not (math.isfinite(k) and k > 0):
        raise ValueError("k must be finite and positive")
    if isinstance(surface, (Sphere, Spheroid, MeridianMesh)):
        method = kw.pop("method", "axisymmetric")
        if method == "axisymmetric":
            n, m_max, ctx = kw.pop("n", None), kw.pop("m_max", None), kw.pop("ctx", None)
            kw.pop("rtol", None)
            if kw:
                raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
            return _bem_axisymmetric(surface, boundary, k, incidence_angle, n=n, m_max=m_max, ctx=ctx)
        if method != "full" or isinstance(surface, MeridianMesh):
            raise ValueError(f"bem(...) supports method=:axisymmetric or :full, got {method}")
        surface = mesh(surface, level=kw.pop("level", 3), ctx=kw.pop("ctx", None),
                       mesh_order=kw.pop("mesh_order", None), qorder=kw.pop("qorder", None))
    else:
        kw.pop("method", None)
    op = Operator(surface, boundary, k, formulation=kw.pop("formulation", None),
                  precision=kw.pop("precision", "double"), refinement=kw.pop("refinement", None),
                  equilibrate=kw.pop("equilibrate", True),
                  _force_promotion=kw.pop("_force_promotion", False))
    for ignored in ("compression", "correction", "gmres_kwargs", "condition_limit"):
        kw.pop(ignored, None)
    if kw:
        raise TypeError(f"unexpected keyword arguments {sorted(kw)}")
    sol = op.solve([(incidence_angle, incidence_azimuth)])
    p, q = sol.traces(0)
    d = sol.diagnostics(0)
    d["unknown_count"] = op.n
    return BEMSolution(surface.body, boundary, float(k), "full", surface, p, q, incidence_angle,
                       incidence_azimuth, d, sol)


def scattering_amplitude(sol, direction=None, *, angle=None, azimuth=None):
    """Far-field amplitude [m] (src/solvers/bem/postprocessing.jl:1-13); default backscatter.
    Axisymmetric solutions take (angle, azimuth) from the axis and the plane
    of incidence (src/solvers/shared/postprocessing.jl:10-14)."""
    if isinstance(sol, AxisymmetricSolution):
        if direction is not None:
            raise ValueError("axisymmetric solutions take angle and azimuth")
        th = math.pi - sol.incidence_angle if angle is None else float(angle)
        ph = math.pi if azimuth is None else float(azimuth)
        return complex(sol.far_field([(th, ph)])[0])
    if direction is None:
        return complex(sol._solution.far_field(0)[0])
    d = np.asarray(direction, dtype=np.float64)
    if d.shape != (3,) or not np.all(np.isfinite(d)):
        raise ValueError("direction must be a finite unit vector with three components")
    return complex(sol._solution.far_field(0, d)[0])


def target_strength(x, **kw):
    """TS = 20 log10 |f| in dB re 1 m^2 (src/postprocessing/target_strength.jl:8,26-28)."""
    if isinstance(x, (BEMSolution, AxisymmetricSolution)):
        x = scattering_amplitude(x, **kw)
    return 20.0 * math.log10(abs(x))


def pressure(solution, points, *, field="total", region=None):
    """pressure(solution, points; field) of a full-3D BEM solution
    (src/fields/pressure.jl:1-20,404-472): one point (3-vector) returns a
    complex number, a 3 x N array (points in columns) or a sequence of
    3-vectors returns an array."""
    if region is not None:
        raise ValueError("region selection requires a coupled fluid BEM solution")
    pts = np.asarray(points, dtype=np.float64)
    single = pts.ndim == 1
    if single and pts.shape != (3,):
        raise ValueError("a point has three coordinates")
    if not single and pts.ndim == 2 and pts.shape[0] != 3 and pts.shape[1] == 3:
        pts = pts.T
    vals = solution._solution.pressure(pts.reshape(3, -1), field)
    return complex(vals[0]) if single else vals


def diagnostics(sol):
    return sol.diagnostics


def incidence_angle_sweep(surface, boundary, k, angles, *, incidence_azimuth=0.0, **kw):
    """Assemble and factor once, solve every incidence (src/postprocessing/sweeps.jl:626-701).
    Returns (amplitudes, target strengths) at backscatter."""
    op = Operator(surface, boundary, k, **kw)
    sol = op.solve([(a, incidence_azimuth) for a in angles])
    amps = np.array([sol.far_field(r)[0] for r in range(len(angles))])
    return amps, 20.0 * np.log10(np.abs(amps))


def frequency_sweep(first, *args, **kw):
    """frequency_sweep(solve, frequencies, c): sequential map over
    k = 2 pi f / c (src/postprocessing/sweeps.jl:103-110).
    frequency_sweep(surface::Mesh, material::FluidFilled, frequencies, c;
    training_frequencies, ...): the reduced-basis sweep (sweeps.jl:208-306),
    reduced_frequency_sweep."""
    if isinstance(first, Mesh):
        return reduced_frequency_sweep(first, *args, **kw)
    solve = first
    frequencies, sound_speed = args
    if not (math.isfinite(sound_speed) and sound_speed > 0):
        raise ValueError("sound_speed must be finite and positive")
    freqs = np.asarray(frequencies, dtype=np.float64)
    if np.any(freqs <= 0):
        raise ValueError("frequencies must be positive")
    ks = 2 * math.pi * freqs / sound_speed
    amps = np.array([scattering_amplitude(solve(k)) for k in ks])
    return {"frequencies": freqs, "k": ks, "ts": 20 * np.log10(np.abs(amps)), "amplitudes": amps}


def reduced_frequency_sweep(surface, material, frequencies, sound_speed, *, training_frequencies,
                            incidence_angle=math.pi / 2, incidence_azimuth=0.0, tolerance=1e-10,
                            rank_tolerance=1e-12, capacity=16, maxbytes=64 * 1024 ** 2, return_diagnostics=False):
    """Reduced-order frequency sweep for dense single-interface fluid CBIE
    (src/postprocessing/sweeps.jl:208-306) on the device: frozen snapshot
    basis from full FP64 training solves; each query accepts the basis
    least-squares solution when its original and row-scaled residuals meet
    `tolerance`, else takes the full solve.  Returns {frequencies, k, ts,
    amplitudes} (backscatter) and, with return_diagnostics, the report."""
    if not isinstance(material, FluidFilled):
        raise ValueError("the reduced sweep solves fluid bodies")
    if not (math.isfinite(sound_speed) and sound_speed > 0):
        raise ValueError("sound_speed must be finite and positive")
    f = np.asarray(frequencies, dtype=np.float64).ravel()
    ft = np.asarray(training_frequencies if training_frequencies is not None else [], dtype=np.float64).ravel()
    for v in (f, ft):
        if v.size == 0 or not np.all(np.isfinite(v)) or np.any(v <= 0):
            raise ValueError("query and training frequencies must be nonempty, finite and positive")
    if not (math.isfinite(incidence_angle) and math.isfinite(incidence_azimuth)):
        raise ValueError("incidence angles must be finite")
    for t in (tolerance, rank_tolerance):
        if not (math.isfinite(t) and 0 < t < 1):
            raise ValueError("residual and rank tolerances must lie in (0, 1)")
    if capacity < 0 or maxbytes < 0:
        raise ValueError("negative storage bound")
    k = 2 * math.pi * f / sound_speed
    kt = 2 * math.pi * ft / sound_speed
    ctx = surface.ctx
    opts = ReducedOptions(float(tolerance), float(rank_tolerance), int(capacity), int(maxbytes))
    amps = np.empty(k.size, dtype=np.complex128)
    tr = (ReducedReport * kt.size)()
    qr = (ReducedReport * k.size)()
    rank = ctypes.c_int32()
    ctx.check(ctx.lib.bem_reduced_sweep(ctx._h, surface._h, material.density_contrast,
                                        material.soundspeed_contrast, kt.size, _d(np.ascontiguousarray(kt)),
                                        k.size, _d(np.ascontiguousarray(k)), float(incidence_angle),
                                        float(incidence_azimuth), ctypes.byref(opts), amps.ctypes.data_as(_dp),
                                        tr, qr, ctypes.byref(rank)))
    sweep = {"frequencies": f, "k": k, "ts": 20 * np.log10(np.abs(amps)), "amplitudes": amps}
    if not return_diagnostics:
        return sweep
    report = {"method": "reduced_basis", "formulation": "cbie", "compression": {"method": "none"},
              "training": [dict(r.as_dict(True), frequency=float(fq)) for r, fq in zip(tr, ft)],
              "queries": [dict(r.as_dict(), frequency=float(fq)) for r, fq in zip(qr, f)],
              "basis_rank": rank.value, "basis_bytes": 16 * 2 * surface.n * rank.value,
              "tolerance": tolerance, "rank_tolerance": rank_tolerance, "capacity": capacity, "maxbytes": maxbytes,
              "incidence_angle": incidence_angle, "incidence_azimuth": incidence_azimuth}
    return {"sweep": sweep, "diagnostics": report}


def debug_lu(A, single=False, b=None, ctx=None, adjoint=False, updates=None, graph=False):
    """Device LU of a dense complex matrix (test hook): returns (LU, 1-based
    ipiv[, x]) with x solving A x = b, or A^H x = b when adjoint.  updates:
    None (the context's shape for n), "split" (bf16-split outer updates,
    single only) or "fp32"; graph: replay the factorization as a captured
    graph."""
    ctx = ctx or default_context()
    if updates not in (None, "split", "fp32") or (updates == "split" and not single):
        raise ValueError("updates must be None, 'fp32' or (single precision) 'split'")
    flags = (1 if single else 0) | (2 if adjoint else 0) | (4 if updates == "split" else 0) | \
        (8 if updates == "fp32" else 0) | (16 if graph else 0)
    A = np.asfortranarray(np.asarray(A, dtype=np.complex128))
    n = A.shape[0]
    LU = np.empty((n, n), dtype=np.complex128, order="F")
    piv = np.empty(n, dtype=np.int64)
    if b is not None:
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.empty(n, dtype=np.complex128)
        ctx.check(ctx.lib.bem_debug_lu(ctx._h, n, A.ctypes.data_as(_dp), flags, LU.ctypes.data_as(_dp),
                                       _i(piv), b.ctypes.data_as(_dp), x.ctypes.data_as(_dp)))
        return LU, piv, x
    ctx.check(ctx.lib.bem_debug_lu(ctx._h, n, A.ctypes.data_as(_dp), flags, LU.ctypes.data_as(_dp),
                                   _i(piv), None, None))
    return LU, piv


def batched_sweep(meshes, problems, angles, ctx=None):
    """Run many independent solves through bem_sweep; returns (amplitudes, latency_ms, diagnostics)."""
    ctx = ctx or meshes[0].ctx
    n = len(meshes)
    hs = (_vp * n)(*[m._h for m in meshes])
    probs = (Problem * n)(*problems)
    ang = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).reshape(n, 2))
    amps = np.empty(n, dtype=np.complex128)
    lat = np.empty(n)
    diags = (Diagnostics * n)()
    ctx.check(ctx.lib.bem_sweep(ctx._h, n, hs, probs, _d(ang), amps.ctypes.data_as(_dp), _d(lat), diags))
    return amps, lat, [d.as_dict() for d in diags]


# ---------------------------------------------------------------------------
# serving
# ---------------------------------------------------------------------------
@dataclass
class ServeResult:
    """One served request: backscatter amplitude, solve report and latencies [ms]."""
    amplitude: complex
    diagnostics: dict
    queue_ms: float
    run_ms: float
    total_ms: float


class Server:
    """Continuous request loop on one GPU (the bem_server_* C ABI).

    `submit` queues the unit of work of bem(surface, boundary, k;
    incidence_angle, incidence_azimuth) followed by scattering_amplitude
    (src/solvers/bem/surface.jl:10-22, src/solvers/bem/postprocessing.jl:1-13)
    and returns a ticket at once; `lanes` requests run concurrently, each
    admitted when its device memory fits the context's HBM budget.  With
    keep_factors, assembled operators and LU factors stay resident and serve
    later requests on the same mesh and problem without reassembly."""

    def __init__(self, ctx=None, lanes=4, keep_factors=True):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        opt = ServerOptions(lanes, 1 if keep_factors else 0)
        h = _vp()
        self.ctx.check(self.lib.bem_server_create(self.ctx._h, ctypes.byref(opt), ctypes.byref(h)))
        self._h = h

    def _check(self, rc, ticket=0):
        if rc != BEM_OK:
            _raise(rc, self.lib.bem_server_error(self._h, ticket).decode())

    def submit(self, surface, boundary, k, *, incidence_angle=math.pi / 2, incidence_azimuth=0.0, **kw):
        """Queue a backscatter request; keyword arguments as for bem()."""
        if not (math.isfinite(k) and k > 0):
            raise ValueError("k must be finite and positive")
        prob = make_problem(boundary, k, formulation=kw.pop("formulation", None),
                            precision=kw.pop("precision", "mixed"), refinement=kw.pop("refinement", None),
                            equilibrate=kw.pop("equilibrate", True))
        if kw:
            raise ValueError(f"unknown options {sorted(kw)}")
        t = ctypes.c_int64()
        self._check(self.lib.bem_server_submit(self._h, surface._h, ctypes.byref(prob), float(incidence_angle),
                                               float(incidence_azimuth), ctypes.byref(t)))
        return t.value

    def poll(self, ticket):
        done = ctypes.c_int32()
        self._check(self.lib.bem_server_poll(self._h, ticket, ctypes.byref(done)))
        return bool(done.value)

    def wait(self, ticket, timeout=None):
        """Block until the request ends; returns a ServeResult or raises its error
        (TimeoutError when `timeout` seconds pass first)."""
        amp = np.empty(2)
        d = Diagnostics()
        tm = np.empty(3)
        rc = self.lib.bem_server_wait(self._h, ticket, -1.0 if timeout is None else 1e3 * timeout, _d(amp),
                                      ctypes.byref(d), _d(tm))
        if rc == BEM_ERR_STATE and timeout is not None and self.poll(ticket) is False:
            raise TimeoutError(f"request {ticket} not finished")
        self._check(rc, ticket)
        return ServeResult(complex(amp[0], amp[1]), d.as_dict(), tm[0], tm[1], tm[2])

    def stats(self):
        st = ServerStats()
        self._check(self.lib.bem_server_stats_get(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self.lib.bem_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# hierarchical-matrix path
# ---------------------------------------------------------------------------
def hmatrix_plan(points, leaf_size=64, eta=2.0, near_radius=0.0):
    """Cluster tree and block tree of the hierarchical path (host only).

    points: 3 x n.  Returns (perm, n_nodes, blocks) with perm the tree order
    (0-based point indices) and blocks an int64 array of rows (lowrank,
    row_lo, m, col_lo, n), dense blocks first."""
    L = load_library()
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).T)
    n = pts.shape[0]
    perm = np.empty(n, dtype=np.int64)
    nn, nd, nl = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    args = (n, _d(pts), int(leaf_size), float(eta), float(near_radius), _i(perm), ctypes.byref(nn),
            ctypes.byref(nd), ctypes.byref(nl))
    if L.bem_hmatrix_plan(*args, None, 0) != BEM_OK:
        raise ValueError("invalid cluster-tree arguments")
    blocks = np.empty((nd.value + nl.value, 5), dtype=np.int64)
    if L.bem_hmatrix_plan(*args, _i(blocks), blocks.size) != BEM_OK:
        raise ValueError("invalid cluster-tree arguments")
    return perm, nn.value, blocks


class HMatrix:
    """Compressed system (the compression=(method=:hmatrix,) path of
    solve_full_bem): rigid surfaces (src/solvers/bem/full.jl:68-184 with
    src/solvers/bem/compression/) and fluid bodies in the Mueller formulation
    (compression/fluid.jl:20-106, one H-matrix over the 2N pressure and
    normal-derivative unknowns).  Cluster tree, dense leaves, ACA blocks in
    complex FP32; solve() runs batched FP64 GMRES for many incidences.  At
    max(k, k_int) * radius <= pi/2 a fluid body's K' and H come from the
    Calderon identities (full.jl:201-207): compressed S and D of each medium
    over one panel tree, S^{-1} by inner GMRES on the compressed S."""

    def __init__(self, surface, boundary, k, *, formulation=None, leaf_size=64, eta=2.0, aca_tolerance=1e-6,
                 max_rank=256, gmres_restart=50, gmres_maxiter=1000, gmres_tolerance=1e-8, preconditioner="twolevel",
                 recycle_dimension=8):
        if not isinstance(boundary, (Rigid, FluidFilled)):
            raise ValueError("the hierarchical path solves rigid surfaces and fluid bodies")
        if not (math.isfinite(k) and k > 0):
            raise ValueError("k must be finite and positive")
        self.mesh = surface
        self.ctx = surface.ctx
        self.k = float(k)
        self.problem = make_problem(boundary, k, formulation=formulation)
        prec = {False: 0, None: 0, "none": 0, True: 1, "block_jacobi": 1, "twolevel": 2}
        if preconditioner not in prec:
            raise ValueError("preconditioner must be :none, :block_jacobi or :twolevel")
        if recycle_dimension < 0:
            raise ValueError("recycle_dimension must be nonnegative")
        o = HMatrixOptions(int(leaf_size), float(eta), float(aca_tolerance), int(max_rank), int(gmres_restart),
                           int(gmres_maxiter), float(gmres_tolerance), prec[preconditioner], int(recycle_dimension))
        h = _vp()
        self.ctx.check(self.ctx.lib.bem_hmatrix_create(self.ctx._h, surface._h, ctypes.byref(self.problem),
                                                       ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        # unknowns: one per panel (rigid), pressure and normal derivative per
        # panel (fluid Mueller: [p; dp/dn], 2N)
        self.n = int(self.info()["n"])

    def __del__(self):
        if getattr(self, "_h", None) and self.ctx._h:
            self.ctx.lib.bem_hmatrix_destroy(self._h)
            self._h = None

    def info(self):
        st = HMatrixInfo()
        self.ctx.check(self.ctx.lib.bem_hmatrix_get_info(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def matvec(self, x, operator="system"):
        """A x for x of shape (n,) or (n, nrhs) in unknown order (panels; for
        fluid bodies the N pressures, then the N normal derivatives).
        operator="rhs" applies the right-hand-side operator of a rigid
        surface instead (b = B q_inc)."""
        which = {"system": 0, "rhs": 1}[operator]
        x = np.asarray(x, dtype=np.complex128)
        one = x.ndim == 1
        if x.shape[0] != self.n or x.ndim > 2:
            raise ValueError(f"x must have {self.n} rows")
        X = np.asfortranarray(x.reshape(self.n, -1))
        Y = np.empty_like(X, order="F")
        self.ctx.check(self.ctx.lib.bem_hmatrix_apply(self._h, which, X.shape[1], X.ctypes.data_as(_dp),
                                                      Y.ctypes.data_as(_dp)))
        return Y[:, 0] if one else Y

    def solve(self, angles):
        """angles: sequence of (incidence_angle, incidence_azimuth)."""
        ang = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).reshape(-1, 2))
        h = _vp()
        self.ctx.check(self.ctx.lib.bem_hmatrix_solve(self._h, ang.shape[0], _d(ang), ctypes.byref(h)))
        return _Solution(self.ctx, h, self.mesh, self.k, ang)


@dataclass(frozen=True)
class Cylinder:
    """Closed cylinder along x, optionally bent (src/geometry/cylinder.jl): radius,
    length of the centerline, radius_curvature (inf: straight) and the depth
    of the ellipsoidal end caps."""
    radius: float
    length: float
    radius_curvature: float = math.inf
    endcap_depth: float = 0.0


def cylinder_arrays(body, n_phi, n_s=None):
    """Mesh of a (bent) capped cylinder: a body of revolution about its axis,
    then the reference's bend map and axis rotation (src/geometry/cylinder.jl:
    15-25,64-70).  Returns nodes 3 x Nv, triangles 3 x N (1-based)."""
    a, L, R, d = body.radius, body.length, body.radius_curvature, body.endcap_depth
    if not (a > 0 and L > 0 and d > 0):
        raise ValueError("need radius, length and endcap_depth > 0")
    if not R > a:
        raise ValueError("radius_curvature must exceed the cylinder radius")
    if math.isfinite(R) and not L / R < 2 * math.pi:
        raise ValueError("the cylinder centerline must span less than a full circle")
    # meridian of the straight template (axis = template z): cap, tube, cap
    th = np.linspace(0.0, np.pi / 2, 4001)
    cap_z, cap_r = -L / 2 - d * np.cos(th), a * np.sin(th)
    tube_z = np.linspace(-L / 2, L / 2, 4001)[1:]
    z = np.concatenate([cap_z, tube_z, (L / 2 + d * np.cos(th[::-1]))[1:]])
    r = np.concatenate([cap_r, np.full(tube_z.size, a), (a * np.sin(th[::-1]))[1:]])
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(z), np.diff(r)))])
    if n_s is None:
        n_s = max(3, int(round(arc[-1] / (2 * np.pi * a / n_phi))) + 1)
    t = np.linspace(0.0, arc[-1], n_s)
    zs, rs = np.interp(t, arc, z), np.interp(t, arc, r)
    rs[0] = rs[-1] = 0.0
    nodes, tri = revolution_arrays(zs, rs, n_phi)
    # template point (x, y, z) = (node y, node z, node x); bend; rotate to x
    x, y, zt = nodes[1], nodes[2], nodes[0]
    if math.isfinite(R):
        s = np.clip(zt, -L / 2, L / 2)
        tt = zt - s
        g = s / R
        bx = 2 * R * np.sin(g / 2) ** 2 + x * np.cos(g) + tt * np.sin(g)
        by = y
        bz = (R - x) * np.sin(g) + tt * np.cos(g)
    else:
        bx, by, bz = x, y, zt
    return np.vstack([bz, bx, by]), tri