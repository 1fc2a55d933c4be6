"""Host-side mirror of the SPECFEM++ Newmark solver interface over the C ABI.

``Engine2D`` / ``Engine3D`` wrap one ``sfx_engine`` (include/specfempp_amd.h).
Their methods carry the names of the reference seams they replace:

=============================  ==================================================
method                          reference
=============================  ==================================================
``initialize()``                ``domain_kernels::initialize(dt)``
                                (include/kokkos_kernels/domain_kernels.hpp:136-163)
``apply_predictor_phase()``     ``newmark::apply_predictor_phase_forward``
                                (include/timescheme/newmark.tpp:140-159)
``update_wavefields(istep)``    ``domain_kernels::update_wavefields<elastic>``
                                (domain_kernels.hpp:69-126)
``apply_corrector_phase()``     ``newmark::apply_corrector_phase_forward``
                                (newmark.tpp:119-137)
``compute_seismograms(isig)``   ``domain_kernels::compute_seismograms``
                                (domain_kernels.hpp:172-186)
``run()``                       ``time_marching<forward>::run``
                                (include/solver/time_marching.tpp:9-124)
=============================  ==================================================

Errors raise :class:`SfxError` with the library's message.  The library is the
HIP build in ``specfempp_amd/_lib``; there is no CPU fallback: without a GPU
the engine constructor fails with the HIP error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _native
from .meshes import material_arrays

DISPLACEMENT, VELOCITY, ACCELERATION = 0, 1, 2
# source time functions (SFX_STF_*)
_STF_CODES = {"ricker": 0, "dirac": 1}
_TYPE_CODES = {"displacement": 0, "velocity": 1, "acceleration": 2, "pressure": 3}


class SfxError(RuntimeError):
    pass


class _Mesh2D(C.Structure):
    _fields_ = [("nspec", C.c_int), ("ngnod", C.c_int), ("npgeo", C.c_int), ("coorg", C.c_void_p),
                ("knods", C.c_void_p), ("material", C.c_void_p), ("nmaterials", C.c_int), ("rho", C.c_void_p),
                ("vp", C.c_void_p), ("vs", C.c_void_p), ("nabs", C.c_int), ("abs_ispec", C.c_void_p),
                ("abs_edge", C.c_void_p), ("nafs", C.c_int), ("afs_ispec", C.c_void_p), ("afs_type", C.c_void_p),
                ("nfs", C.c_int), ("fs_acoustic", C.c_void_p), ("fs_elastic", C.c_void_p)]


class _Mesh3D(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("x0", C.c_double), ("y0", C.c_double),
                ("z0", C.c_double), ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
                ("material", C.c_void_p), ("nmaterials", C.c_int), ("rho", C.c_void_p), ("vp", C.c_void_p),
                ("vs", C.c_void_p), ("absorbing", C.c_int * 6)]


class _Mesh3DDB(C.Structure):
    _fields_ = [("nspec", C.c_int), ("nglob", C.c_int), ("ibool", C.c_void_p), ("x", C.c_void_p),
                ("y", C.c_void_p), ("z", C.c_void_p), ("nspec_irregular", C.c_int),
                ("irregular_element_number", C.c_void_p), ("xix_regular", C.c_float),
                ("jacobian_regular", C.c_float)] + \
               [(n, C.c_void_p) for n in ("xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz",
                                          "jacobian", "kappa", "mu", "rho", "rho_vp", "rho_vs", "ispec_type")] + \
               [("nabs", C.c_int), ("abs_ispec", C.c_void_p), ("abs_ijk", C.c_void_p),
                ("abs_jacobian2Dw", C.c_void_p), ("abs_normal", C.c_void_p)]


class _Source(C.Structure):
    _fields_ = [("kind", C.c_int), ("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("angle", C.c_double),
                ("fx", C.c_double), ("fy", C.c_double), ("fz", C.c_double), ("m", C.c_double * 6),
                ("f0", C.c_double), ("tshift", C.c_double), ("factor", C.c_double), ("stf", C.c_int)]


class _Receiver(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double), ("angle", C.c_double)]


class _Run(C.Structure):
    _fields_ = [("dt", C.c_double), ("nstep", C.c_int), ("nstep_between_samples", C.c_int),
                ("nseismo_types", C.c_int), ("seismo_types", C.c_int * 4), ("chunk_size", C.c_int),
                ("scatter", C.c_int), ("start_time", C.c_double)]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _sources(sources, dim):
    arr = (_Source * max(len(sources), 1))()
    for i, s in enumerate(sources):
        kind = s.get("type", "force")
        arr[i].kind = 1 if kind == "moment_tensor" else 0
        arr[i].x = s["x"]
        arr[i].y = s.get("y", 0.0)
        arr[i].z = s["z"]
        arr[i].angle = s.get("angle", 0.0)
        arr[i].fx, arr[i].fy, arr[i].fz = s.get("fx", 0.0), s.get("fy", 0.0), s.get("fz", 0.0)
        if dim == 2:
            m = [s.get("mxx", 0.0), s.get("mzz", 0.0), s.get("mxz", 0.0), 0.0, 0.0, 0.0]
        else:
            m = [s.get("mxx", 0.0), s.get("myy", 0.0), s.get("mzz", 0.0), s.get("mxy", 0.0), s.get("mxz", 0.0),
                 s.get("myz", 0.0)]
        for k in range(6):
            arr[i].m[k] = m[k]
        arr[i].stf = _STF_CODES.get(s.get("stf") or "ricker", -1)  # the library refuses unknown ones
        arr[i].f0 = s.get("f0", 0.0) if arr[i].stf == _STF_CODES["dirac"] else s["f0"]
        arr[i].tshift = s.get("tshift", 0.0)
        arr[i].factor = s.get("factor", 1.0)
    return arr


def _receivers(receivers):
    arr = (_Receiver * max(len(receivers), 1))()
    for i, r in enumerate(receivers):
        arr[i].x, arr[i].y, arr[i].z, arr[i].angle = r["x"], r.get("y", 0.0), r["z"], r.get("angle", 0.0)
    return arr


# specfem::element::medium_tag, specfem::wavefield::simulation_field and the
# Newmark phase direction, as the library's SFX_MEDIUM_*, SFX_WAVEFIELD_* and
# SFX_DIRECTION_* codes
MEDIA = {"elastic_psv": 0, "acoustic": 1, "elastic_sh": 2, "poroelastic": 3, "elastic_psv_t": 4}
WAVEFIELDS = {"forward": 0, "adjoint": 1, "backward": 2, "buffer": 3}
DIRECTIONS = {"forward": 0, "backward": 1}

SCATTER_COLORED, SCATTER_ATOMIC, SCATTER_GATHER, SCATTER_BLOCK = 0, 1, 2, 3
_SCATTER_CODES = {"colored": 0, "atomic": 1, "gather": 2, "block": 3}


def _run_desc(dt, nstep, nsb, types, chunk, scatter=SCATTER_COLORED, start_time=0.0):
    run = _Run()
    run.dt, run.nstep, run.nstep_between_samples, run.chunk_size = dt, nstep, nsb, chunk
    run.start_time = float(start_time)
    run.scatter = _SCATTER_CODES.get(scatter, scatter)
    if len(types) > 4:
        raise ValueError(f"at most 4 seismogram types, got {len(types)}")
    run.nseismo_types = len(types)
    for i, t in enumerate(types):
        run.seismo_types[i] = _TYPE_CODES[t] if isinstance(t, str) else int(t)
    return run


class _EngineBase:
    dim = 0

    def _check(self, rc):
        if rc < 0:
            raise SfxError(self.lib.sfx_last_error().decode())
        return rc

    def _finish_init(self, types):
        self.types = list(types)
        sizes = (C.c_int64 * 7)()
        self._check(self.lib.sfx_get_sizes(self.h, sizes))
        self.dim, self.nspec, self.nglob, self.ngll, self.nsig, self.nrec, self.nsrc = [int(x) for x in sizes]
        self.ncomp = self.dim
        t0 = C.c_double()
        self._check(self.lib.sfx_get_t0(self.h, C.byref(t0)))
        self.t0 = t0.value
        self.field_nglob = self.nglob  # points of the (elastic) field

    def close(self):
        if getattr(self, "h", None):
            self.lib.sfx_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- data contract ----
    def index_mapping(self):
        npe = 25 if self.dim == 2 else 125
        out = np.zeros(self.nspec * npe, dtype=np.int32)
        self._check(self.lib.sfx_get_index_mapping(self.h, _ptr(out)))
        return out.reshape((self.nspec,) + (5,) * self.dim)

    def field_index(self):
        out = np.zeros(self.nglob, dtype=np.int32)
        self._check(self.lib.sfx_get_field_index(self.h, _ptr(out)))
        return out

    def compute_to_mesh(self):
        out = np.zeros(self.nspec, dtype=np.int32)
        self._check(self.lib.sfx_get_compute_to_mesh(self.h, _ptr(out)))
        return out

    def locations(self, which="sources"):
        n = self.nsrc if which == "sources" else self.nrec
        ispec = np.zeros(n, dtype=np.int32)
        loc = np.zeros((n, self.dim))
        self._check(self.lib.sfx_get_locations(self.h, 0 if which == "sources" else 1, _ptr(ispec), _ptr(loc)))
        return ispec, loc

    # ---- time scheme seam ----
    def initialize(self):
        self._check(self.lib.sfx_initialize(self.h))

    def apply_predictor_phase(self):
        return self._check(self.lib.sfx_predictor(self.h))

    def update_wavefields(self, istep):
        return self._check(self.lib.sfx_update_wavefields(self.h, istep))

    def apply_corrector_phase(self):
        return self._check(self.lib.sfx_corrector(self.h))

    def compute_seismograms(self, isig):
        self._check(self.lib.sfx_compute_seismograms(self.h, isig))

    # ---- wavefield / medium / direction seams ----
    def initialize_wavefield(self, wavefield="forward"):
        self._check(self.lib.sfx_initialize_wavefield(self.h, WAVEFIELDS[wavefield]))

    def predictor_phase(self, medium="elastic_psv", wavefield="forward", direction="forward"):
        return self._check(self.lib.sfx_predictor_phase(self.h, MEDIA[medium], WAVEFIELDS[wavefield],
                                                        DIRECTIONS[direction]))

    def corrector_phase(self, medium="elastic_psv", wavefield="forward", direction="forward"):
        return self._check(self.lib.sfx_corrector_phase(self.h, MEDIA[medium], WAVEFIELDS[wavefield],
                                                        DIRECTIONS[direction]))

    def update_wavefields_medium(self, medium, wavefield, istep):
        return self._check(self.lib.sfx_update_wavefields_medium(self.h, MEDIA[medium], WAVEFIELDS[wavefield],
                                                                 istep))

    def compute_seismograms_wavefield(self, wavefield, isig):
        self._check(self.lib.sfx_compute_seismograms_wavefield(self.h, WAVEFIELDS[wavefield], isig))

    def copy_wavefield(self, dst, src):
        self._check(self.lib.sfx_copy_wavefield(self.h, WAVEFIELDS[dst], WAVEFIELDS[src]))

    def compute_derivatives(self, dt):
        self._check(self.lib.sfx_compute_derivatives(self.h, float(dt)))

    def wavefield_field(self, wavefield, which):
        codes = {"displacement": 0, "velocity": 1, "acceleration": 2, "mass_inverse": 3}
        out = np.zeros((self.ncomp, self.field_nglob), dtype=np.float32)
        self._check(self.lib.sfx_get_wavefield(self.h, WAVEFIELDS[wavefield], codes[which], _ptr(out)))
        return out

    def set_wavefield_field(self, wavefield, which, values):
        codes = {"displacement": 0, "velocity": 1, "acceleration": 2}
        arr = np.ascontiguousarray(values, dtype=np.float32).reshape(self.ncomp, self.field_nglob)
        self._check(self.lib.sfx_set_wavefield(self.h, WAVEFIELDS[wavefield], codes[which], _ptr(arr)))

    # ---- fused path ----
    def step(self, istep):
        self._check(self.lib.sfx_step(self.h, istep))

    def run(self, first=0, count=None):
        if count is None:
            count = self.nstep - first
        self._check(self.lib.sfx_run(self.h, first, count))

    def run_noalloc(self, first=0, count=None):
        """``run`` without device allocation (``sfx_run_noalloc``): 3-D gather
        engines of at most ``SFX_NOALLOC_MAX_ELEMENTS`` elements; the element
        forces live in the library's static per-GPU scratch."""
        if count is None:
            count = self.nstep - first
        self._check(self.lib.sfx_run_noalloc(self.h, first, count))

    def synchronize(self):
        self._check(self.lib.sfx_synchronize(self.h))

    def layout(self):
        """The storage the engine chose: ``geometry`` "point", "element" or
        "control_nodes_4" / "control_nodes_9" (recomputed per point from the
        element's control nodes, the same bits as the per-point arrays),
        ``element_properties`` (kappa and mu read once per element),
        ``scatter`` and ``row_order`` (points in coordinate row order)."""
        out = (C.c_int64 * 4)()
        self._check(self.lib.sfx_get_layout(self.h, out))
        geom = {0: "point", 1: "element", 4: "control_nodes_4", 9: "control_nodes_9"}[int(out[0])]
        scat = {0: "colored", 1: "atomic", 2: "gather", 3: "block"}.get(int(out[2]), str(int(out[2])))
        return {"geometry": geom, "element_properties": bool(out[1]), "scatter": scat, "row_order": bool(out[3])}

    def timings(self):
        out = (C.c_double * 4)()
        self._check(self.lib.sfx_get_timings(self.h, out))
        return {"stiffness_ms": out[0], "update_ms": out[1], "boundary_source_seismo_ms": out[2], "step_ms": out[3]}

    # ---- outputs ----
    def seismograms(self):
        out = np.zeros((self.nsig, len(self.types), self.nrec, self.ncomp), dtype=np.float32)
        self._check(self.lib.sfx_get_seismograms(self.h, _ptr(out)))
        return out

    def field(self, which):
        codes = {"displacement": 0, "velocity": 1, "acceleration": 2, "mass_inverse": 3}
        out = np.zeros((self.ncomp, self.field_nglob), dtype=np.float32)
        self._check(self.lib.sfx_get_field(self.h, codes[which], _ptr(out)))
        return out

    def set_field(self, which, values):
        codes = {"displacement": 0, "velocity": 1, "acceleration": 2}
        arr = np.ascontiguousarray(values, dtype=np.float32).reshape(self.ncomp, self.field_nglob)
        self._check(self.lib.sfx_set_field(self.h, codes[which], _ptr(arr)))


class Engine2D(_EngineBase):
    """2-D elastic P-SV engine built from a :class:`Mesh2D`-like object."""

    KERNEL_NAMES = ("rho", "mu", "kappa", "rhop", "alpha", "beta")

    # ---- adjoint (time_marching<combined>) ----
    def forward_record(self, snapshot_every=0):
        """Whole forward loop that keeps the Stacey-element record and the
        final state for :meth:`adjoint_run`; snapshot_every > 0 also keeps
        checkpoints the backward reconstruction is reset to."""
        self._check(self.lib.sfx_forward_record(self.h, int(snapshot_every)))

    def set_adjoint_sources(self, positions, stf):
        """positions: [{"x", "z"}]; stf (nadj, nstep, 2) in forward time order."""
        arr = np.ascontiguousarray(stf, dtype=np.float32)
        if arr.shape != (len(positions), self.nstep, 2):
            raise ValueError(f"stf must be (nadj, nstep, 2) = ({len(positions)}, {self.nstep}, 2)")
        self._adj_keep = (arr, _receivers(positions))
        self._check(self.lib.sfx_set_adjoint_sources(self.h, len(positions), self._adj_keep[1], _ptr(arr)))

    def set_properties(self, rho, kappa, mu):
        """Per-point rho, kappa (bulk modulus), mu, each (nspec, 5, 5) in mesh element order."""
        arrs = [np.ascontiguousarray(x, dtype=np.float32).reshape(self.nspec, 5, 5) for x in (rho, kappa, mu)]
        self._check(self.lib.sfx_set_properties(self.h, *[_ptr(x) for x in arrs]))

    def adjoint_run(self):
        """Combined adjoint + backward loop; seismograms() then returns the
        backward field's, latest sample first."""
        self._check(self.lib.sfx_adjoint_run(self.h))

    def stacey_elements(self):
        """Number of elements with an absorbing edge (the Stacey elements,
        last in compute order)."""
        return self._nstacey

    def point_coordinates(self):
        """GLL point coordinates (nspec, 5, 5, 2) in compute order (float)."""
        out = np.zeros((self.nspec, 5, 5, 2), dtype=np.float32)
        self._check(self.lib.sfx_get_point_coordinates(self.h, _ptr(out)))
        return out

    def stacey_record(self):
        """The last forward_record's Stacey record, (nstacey, 5, 5, nstep, 2)
        in the reference's LayoutLeft boundary-values layout."""
        out = np.zeros((self._nstacey, 5, 5, self.nstep, 2), dtype=np.float32, order="F")
        self._check(self.lib.sfx_get_stacey_record(self.h, _ptr(out)))
        return out

    def set_stacey_record(self, record):
        """Load a Stacey record (as ``stacey_record`` returns it); the forward
        wavefield is then taken to hold the recorded run's final state."""
        arr = np.asfortranarray(record, dtype=np.float32)
        if arr.shape != (self._nstacey, 5, 5, self.nstep, 2):
            raise ValueError(f"record must be (nstacey, 5, 5, nstep, 2) = ({self._nstacey}, 5, 5, {self.nstep}, 2)")
        self._check(self.lib.sfx_set_stacey_record(self.h, _ptr(arr)))

    def point_weights(self):
        """w_i w_j J of every element point, (nspec, 5, 5) in mesh element order."""
        out = np.zeros((self.nspec, 5, 5), dtype=np.float32)
        self._check(self.lib.sfx_get_point_weights(self.h, _ptr(out)))
        return out

    def kernels(self):
        """dict name -> (nspec, 5, 5) in mesh element order."""
        out = np.zeros((6, self.nspec, 5, 5), dtype=np.float32)
        self._check(self.lib.sfx_get_kernels(self.h, _ptr(out)))
        return dict(zip(self.KERNEL_NAMES, out))

    def __init__(self, mesh, sources, receivers, dt, nstep, nstep_between_samples=1,
                 seismogram_types=("displacement",), chunk_size=1, scatter="colored", host_only=False,
                 start_time=0.0):
        """host_only: build the host assembly only (numbering, locations; usable
        by the data-contract getters and ``Batch``), no device memory.
        start_time: 0 for the automatic start (earliest source onset), else the
        user start time (the configuration's time-scheme t0, negated)."""
        self.lib = _native.load()
        self.host_only = host_only
        self.nstep = nstep
        self.dt = dt
        coorg = np.ascontiguousarray(np.asarray(mesh.coorg, dtype=np.float64).T)
        knods = np.ascontiguousarray(np.asarray(mesh.knods, dtype=np.int32).T)
        kmato = np.ascontiguousarray(mesh.kmato, dtype=np.int32)
        rho, vp, vs = material_arrays(mesh.materials)
        abs_ispec = np.ascontiguousarray(mesh.abs_ispec, dtype=np.int32)
        abs_edge = np.ascontiguousarray(mesh.abs_type, dtype=np.int32)

        def ints(name):
            v = getattr(mesh, name, None)
            return np.ascontiguousarray(v if v is not None else [], dtype=np.int32)

        afs_ispec, afs_type = ints("afs_ispec"), ints("afs_type")
        fs_ac, fs_el = ints("fs_acoustic"), ints("fs_elastic")
        acoustic_mat = vs == 0.0
        self._nstacey = int(np.unique(abs_ispec[~acoustic_mat[kmato[abs_ispec]]]).size) if len(abs_ispec) else 0
        self._keep = (coorg, knods, kmato, rho, vp, vs, abs_ispec, abs_edge, afs_ispec, afs_type, fs_ac, fs_el)
        m = _Mesh2D(mesh.nspec, mesh.ngnod, mesh.npgeo, _ptr(coorg), _ptr(knods), _ptr(kmato), len(rho), _ptr(rho),
                    _ptr(vp), _ptr(vs), len(abs_ispec), _ptr(abs_ispec), _ptr(abs_edge), len(afs_ispec),
                    _ptr(afs_ispec), _ptr(afs_type), len(fs_ac), _ptr(fs_ac), _ptr(fs_el))
        run = _run_desc(dt, nstep, nstep_between_samples, seismogram_types, chunk_size, scatter, start_time)
        h = C.c_void_p()
        self.h = None
        create = self.lib.sfx_engine_create_host_2d if host_only else self.lib.sfx_engine_create_2d
        self._check(create(C.byref(m), len(sources), _sources(sources, 2), len(receivers), _receivers(receivers),
                           C.byref(run), C.byref(h)))
        self.h = h
        self._finish_init(seismogram_types)
        sizes = (C.c_int64 * 4)()
        self._check(self.lib.sfx_get_media_sizes(self.h, sizes))
        self.nspec_elastic, self.nspec_acoustic, self.nglob_elastic, self.nglob_acoustic = [int(x) for x in sizes]
        self.field_nglob = self.nglob_elastic

    # ---- acoustic medium ----
    def field_index_acoustic(self):
        """Acoustic field index of every global point (-1 off the acoustic elements)."""
        out = np.zeros(self.nglob, dtype=np.int32)
        self._check(self.lib.sfx_get_field_index_acoustic(self.h, _ptr(out)))
        return out

    def acoustic_field(self, which):
        """chi ("displacement"), chi_dot, chi_dot_dot or the mass inverse of the
        acoustic field, (nglob_acoustic,) on the acoustic field index."""
        codes = {"displacement": 0, "velocity": 1, "acceleration": 2, "mass_inverse": 3}
        out = np.zeros(self.nglob_acoustic, dtype=np.float32)
        self._check(self.lib.sfx_get_medium_field(self.h, MEDIA["acoustic"], codes[which], _ptr(out)))
        return out


class Engine3D(_EngineBase):
    """3-D elastic engine on a structured hexahedral mesh (see ``mesh3d``).

    ``scatter="auto"`` picks the fastest measured path for the mesh (see
    ``AUTO_3D``); "block", "gather", "colored" and "atomic" force one.
    """

    # path "auto" selects: "gather" (block assembly needs even nx, ny, nz and
    # is only chosen where it is faster)
    AUTO_3D = "gather"

    def __init__(self, mesh, sources, receivers, dt, nstep, nstep_between_samples=1,
                 seismogram_types=("displacement",), scatter="auto", host_only=False, start_time=0.0):
        if scatter == "auto":
            even = all(int(mesh[k]) % 2 == 0 for k in ("nx", "ny", "nz"))
            scatter = self.AUTO_3D if (self.AUTO_3D != "block" or even) else "gather"
        self.scatter = scatter
        self.host_only = host_only
        self.lib = _native.load()
        self.nstep = nstep
        self.dt = dt
        mat = np.ascontiguousarray(mesh["material"], dtype=np.int32).ravel()
        rho = np.ascontiguousarray(mesh["rho"], dtype=np.float64)
        vp = np.ascontiguousarray(mesh["vp"], dtype=np.float64)
        vs = np.ascontiguousarray(mesh["vs"], dtype=np.float64)
        self._keep = (mat, rho, vp, vs)
        m = _Mesh3D()
        m.nx, m.ny, m.nz = mesh["nx"], mesh["ny"], mesh["nz"]
        m.x0, m.y0, m.z0 = mesh.get("x0", 0.0), mesh.get("y0", 0.0), mesh.get("z0", 0.0)
        m.hx, m.hy, m.hz = mesh["hx"], mesh["hy"], mesh["hz"]
        m.material, m.nmaterials = _ptr(mat), len(rho)
        m.rho, m.vp, m.vs = _ptr(rho), _ptr(vp), _ptr(vs)
        for k in range(6):
            m.absorbing[k] = int(mesh["absorbing"][k])
        run = _run_desc(dt, nstep, nstep_between_samples, seismogram_types, 1, scatter, start_time)
        h = C.c_void_p()
        self.h = None
        create = self.lib.sfx_engine_create_host_3d if host_only else self.lib.sfx_engine_create_3d
        self._check(create(C.byref(m), len(sources), _sources(sources, 3), len(receivers), _receivers(receivers),
                           C.byref(run), C.byref(h)))
        self.h = h
        self._finish_init(seismogram_types)


class Engine3DDatabase(_EngineBase):
    """3-D elastic engine on a mesh database (general hexahedra; see
    ``specfempp_amd.io.database3d``): the reference's 3-D mesh as
    read_3d_mesh holds it, with per-point geometry on irregular elements.
    ``scatter`` is "colored" (greedy element coloring) or "atomic"."""

    def __init__(self, db, sources, receivers, dt, nstep, nstep_between_samples=1,
                 seismogram_types=("displacement",), scatter="colored", host_only=False, start_time=0.0):
        self.lib = _native.load()
        self.host_only = host_only
        self.nstep, self.dt, self.scatter = nstep, dt, scatter
        nspec = db.nspec
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32).reshape(-1)  # noqa: E731
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).reshape(-1)  # noqa: E731
        ab = db.absorbing
        arrays = dict(ibool=i32(db.ibool), x=f32(db.x), y=f32(db.y), z=f32(db.z),
                      irregular_element_number=i32(db.irregular_element_number),
                      kappa=f32(db.kappa), mu=f32(db.mu), rho=f32(db.rho), rho_vp=f32(db.rho_vp),
                      rho_vs=f32(db.rho_vs), ispec_type=i32(db.ispec_type), abs_ispec=i32(ab.ispec),
                      abs_ijk=i32(ab.ijk), abs_jacobian2Dw=f32(ab.jacobian2Dw), abs_normal=f32(ab.normal))
        for n in ("xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz", "jacobian"):
            arrays[n] = f32(db.derivs[n])
        self._keep = arrays
        m = _Mesh3DDB()
        m.nspec, m.nglob = nspec, db.nglob
        m.nspec_irregular = len(db.irregular_element_number)
        m.xix_regular, m.jacobian_regular = float(db.xix_regular), float(db.jacobian_regular)
        m.nabs = len(ab.ispec)
        for k, v in arrays.items():
            setattr(m, k, _ptr(v))
        run = _run_desc(dt, nstep, nstep_between_samples, seismogram_types, 1, scatter, start_time)
        h = C.c_void_p()
        self.h = None
        create = self.lib.sfx_engine_create_host_3d_db if host_only else self.lib.sfx_engine_create_3d_db
        self._check(create(C.byref(m), len(sources), _sources(sources, 3), len(receivers), _receivers(receivers),
                           C.byref(run), C.byref(h)))
        self.h = h
        self._finish_init(seismogram_types)


class Batch:
    """Many 2-D engines advanced together (``sfx_batch_*``).

    The engines are only read at creation; the batch holds its own copy of
    their device data, runs every simulation to its own step count from a zero
    state, and returns each simulation's seismograms laid out as its engine's.
    """

    def __init__(self, engines):
        self.lib = _native.load()
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h for e in self.engines])
        h = C.c_void_p()
        self.h = None
        self._check(self.lib.sfx_batch_create(len(self.engines), arr, C.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc < 0:
            raise SfxError(self.lib.sfx_last_error().decode())
        return rc

    _SCATTERS = {"gather": 2, "colored": 0, "atomic": 1}

    def tune(self, steps_per_candidate=10):
        """Time every scatter, elements per workgroup and graph use on this
        batch and keep the fastest."""
        out = (C.c_int * 4)()
        self._check(self.lib.sfx_batch_tune(self.h, steps_per_candidate, out))
        names = {v: k for k, v in self._SCATTERS.items()}
        return {"elems_per_block": out[0], "graphs": bool(out[1]), "colors": out[2], "scatter": names[out[3]]}

    def set_config(self, elems_per_block=5, graphs=True, scatter="gather"):
        """scatter "gather": one launch per time step; "colored" / "atomic": one
        launch per kernel family per step."""
        if scatter not in self._SCATTERS:
            raise ValueError(f"unknown batch scatter {scatter!r}")
        self._check(self.lib.sfx_batch_set_config(self.h, self._SCATTERS[scatter], elems_per_block, int(graphs)))

    def run(self):
        self._check(self.lib.sfx_batch_run(self.h))

    def seismograms(self, i):
        e = self.engines[i]
        out = np.zeros((e.nsig, len(e.types), e.nrec, e.ncomp), dtype=np.float32)
        self._check(self.lib.sfx_batch_get_seismograms(self.h, i, _ptr(out)))
        return out

    def info(self):
        out = (C.c_double * 8)()
        self._check(self.lib.sfx_batch_get_info(self.h, out))
        names = {v: k for k, v in self._SCATTERS.items()}
        return {"run_ms": out[0], "elems_per_block": int(out[1]), "graphs": bool(out[2]), "colors": int(out[3]),
                "padded_points": int(out[4]), "point_steps": out[5], "scatter": names[int(out[6])],
                "launches_per_step": int(out[7])}

    def close(self):
        if getattr(self, "h", None):
            self.lib.sfx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
