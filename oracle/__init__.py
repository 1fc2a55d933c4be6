"""ORACLE — test infrastructure only.

ctypes binding of the CPU restatement in ``oracle/*.cpp``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
this module, and only as the checker: the product library never loads it.

The library is built by ``oracle/Makefile`` into ``oracle/_build/liboracle.so``
(``__graft_entry__.build()`` runs the recipe).  If it is missing, ``load()``
builds it once with ``make``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from specfempp_amd.meshes import material_arrays

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    lib.orc_last_error.restype = C.c_char_p
    vp = C.c_void_p
    lib.orc2d_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip, C.c_int, _dp, _dp, _dp,
                                 C.c_int, _ip, _ip, C.c_int, _ip, _ip, C.c_int, _ip, _ip, C.c_int, C.c_double,
                                 C.POINTER(vp)]
    lib.orc2d_nglob_media.argtypes = [vp, _ip]
    lib.orc2d_acoustic.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.orc2d_destroy.argtypes = [vp]
    lib.orc2d_set_start_time.argtypes = [vp, C.c_double]
    lib.orc2d_nglob.argtypes = [vp]
    lib.orc2d_numbering.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.orc2d_geometry.argtypes = [vp] + [C.c_void_p] * 6
    lib.orc2d_mass_inverse.argtypes = [vp, _dp]
    lib.orc2d_quadrature.argtypes = [vp, _dp, _dp, _dp]
    lib.orc2d_locate.argtypes = [vp, C.c_double, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
    lib.orc2d_forward.argtypes = [vp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_double, C.c_int, C.c_int, _ip,
                                  _dp, C.POINTER(C.c_double), C.c_int, C.c_void_p]
    lib.orc2d_apply_stiffness.argtypes = [vp, _dp, _dp, _dp]
    lib.orc2d_nstacey.argtypes = [vp]
    lib.orc2d_sample.argtypes = [vp, C.c_int, _dp, C.c_int, _ip, _dp, _dp]
    lib.orc2d_set_properties.argtypes = [vp, _dp, _dp, _dp]
    lib.orc2d_properties.argtypes = [vp, _dp, _dp, _dp, _dp]
    lib.orc2d_forward_record.argtypes = [vp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_double, C.c_int, C.c_int,
                                         _ip, _dp, C.POINTER(C.c_double), C.c_int, _dp, _dp, C.c_int, _dp]
    lib.orc2d_combined.argtypes = [vp, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, _dp, C.c_int, C.c_double,
                                   C.c_int, C.c_int, _ip, _dp, _dp, _dp, _dp, C.c_int, C.c_int, _dp]
    lib.orc3d_last_error.restype = C.c_char_p
    lib.orc3d_create.argtypes = [C.c_int] + [C.c_int] * 3 + [C.c_double] * 6 + [_ip, _dp, _dp, _dp, _ip, C.c_double,
                                                                                 C.POINTER(vp)]
    fp = C.c_void_p
    lib.orc3d_create_db.argtypes = ([C.c_int, C.c_int, C.c_int, _ip, fp, fp, fp, C.c_int, _ip, C.c_float, C.c_float,
                                     C.POINTER(C.c_void_p)] + [fp] * 6 + [C.c_int, _ip, _ip, fp, fp, C.c_double,
                                                                          C.POINTER(vp)])
    lib.orc3d_destroy.argtypes = [vp]
    lib.orc3d_nglob.argtypes = [vp]
    lib.orc3d_ibool.argtypes = [vp, _ip]
    lib.orc3d_mass_inverse.argtypes = [vp, _dp]
    lib.orc3d_forward.argtypes = [vp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_double, C.c_int, C.c_int, _ip, _dp,
                                  C.POINTER(C.c_double), C.c_void_p]
    lib.orc3d_apply_stiffness.argtypes = [vp, _dp, _dp]
    lib.orc3d_set_threads.argtypes = [C.c_int]
    lib.orc3d_set_threads.restype = None
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise RuntimeError(load().orc_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


SEISMO_TYPES = {"displacement": 0, "velocity": 1, "acceleration": 2, "pressure": 3}


class Oracle2D:
    """Assembly + forward solver of the restatement for one mesh.

    ``mesh`` is a :class:`specfempp_amd.io.fortran_db.Mesh2D` or any object
    with the same fields (``nspec, ngnod, npgeo, coorg (2, npgeo), knods
    (ngnod, nspec), kmato, materials[i].{density,cp,cs}, abs_ispec,
    abs_type``).  ``precision`` is 4 (type_real = float) or 8 (double).
    """

    def __init__(self, mesh, precision: int = 4, chunk: int = 1, dt: float = 0.0):
        lib = load()
        self.mesh = mesh
        self.precision = precision
        coorg = np.ascontiguousarray(np.asarray(mesh.coorg, dtype=np.float64).T)
        knods = np.ascontiguousarray(np.asarray(mesh.knods, dtype=np.int32).T)
        kmato = np.ascontiguousarray(mesh.kmato, dtype=np.int32)
        rho, cp, cs = material_arrays(mesh.materials)
        abs_ispec = np.ascontiguousarray(mesh.abs_ispec, dtype=np.int32)
        abs_type = np.ascontiguousarray(mesh.abs_type, dtype=np.int32)

        def arr(name):
            v = getattr(mesh, name, None)
            return np.ascontiguousarray(np.zeros(0) if v is None else v, dtype=np.int32)

        afs_ispec, afs_type = arr("afs_ispec"), arr("afs_type")
        fs_ac, fs_el = arr("fs_acoustic"), arr("fs_elastic")
        h = C.c_void_p()
        _check(lib.orc2d_create(precision, mesh.nspec, mesh.ngnod, mesh.npgeo, coorg, knods, kmato, len(rho), rho, cp,
                                cs, len(abs_ispec), abs_ispec, abs_type, len(afs_ispec), afs_ispec, afs_type,
                                len(fs_ac), fs_ac, fs_el, chunk, dt, C.byref(h)))
        self.h = h
        self.nspec = mesh.nspec
        self.nglob = lib.orc2d_nglob(h)
        n2 = np.zeros(2, dtype=np.int32)
        lib.orc2d_nglob_media(h, n2)
        self.nglob_el, self.nglob_ac = int(n2[0]), int(n2[1])

    def __del__(self):
        if getattr(self, "h", None):
            load().orc2d_destroy(self.h)
            self.h = None

    def numbering(self):
        ibool = np.zeros((self.nspec, 5, 5), dtype=np.int32)
        midx = np.zeros(self.nglob, dtype=np.int32)
        c2m = np.zeros(self.nspec, dtype=np.int32)
        load().orc2d_numbering(self.h, _ptr(ibool), _ptr(midx), _ptr(c2m))
        return ibool, midx, c2m

    def geometry(self):
        coord = np.zeros((self.nspec, 5, 5, 2))
        arrs = [np.zeros((self.nspec, 5, 5)) for _ in range(5)]
        load().orc2d_geometry(self.h, _ptr(coord), *[_ptr(a) for a in arrs])
        return coord, arrs  # xix, xiz, gammax, gammaz, jacobian

    def mass_inverse(self):
        out = np.zeros((2, self.nglob_el))
        load().orc2d_mass_inverse(self.h, out)
        return out

    def acoustic(self):
        """(acoustic field index of every iglob (-1 elsewhere), medium of every
        compute element (0 elastic, 1 acoustic), acoustic mass inverse)."""
        idx = np.zeros(self.nglob, dtype=np.int32)
        med = np.zeros(self.nspec, dtype=np.int32)
        minv = np.zeros(self.nglob_ac)
        load().orc2d_acoustic(self.h, _ptr(idx), _ptr(med), _ptr(minv))
        return idx, med, minv

    def quadrature(self):
        xi, w, hp = np.zeros(5), np.zeros(5), np.zeros(25)
        load().orc2d_quadrature(self.h, xi, w, hp)
        return xi, w, hp.reshape(5, 5)

    def locate(self, x: float, z: float):
        i, xi, g = C.c_int(), C.c_double(), C.c_double()
        _check(load().orc2d_locate(self.h, x, z, C.byref(i), C.byref(xi), C.byref(g)))
        return i.value, xi.value, g.value

    @staticmethod
    def _src_rows(sources):
        src = np.zeros((len(sources), 11))
        for i, s in enumerate(sources):
            kind = 1 if s.get("type", "force") == "moment_tensor" else 0
            stf = 1 if s.get("stf", "ricker") == "dirac" else 0
            src[i] = [kind, s["x"], s["z"], s.get("angle", 0.0), s.get("mxx", 0.0), s.get("mzz", 0.0),
                      s.get("mxz", 0.0), s.get("f0", 0.0), s.get("tshift", 0.0), s["factor"], stf]
        return src

    @staticmethod
    def _rec_rows(receivers):
        rec = np.zeros((len(receivers), 3))
        for i, r in enumerate(receivers):
            rec[i] = [r["x"], r["z"], r.get("angle", 0.0)]
        return rec

    def sample(self, receivers, fields, types=("displacement",)):
        """Seismogram values (ntypes, nrec, 2) of given fields (3, 2, nglob) [u, v, a]."""
        rec = self._rec_rows(receivers)
        tcodes = np.array([SEISMO_TYPES[t] for t in types], dtype=np.int32)
        out = np.zeros((len(types), len(receivers), 2))
        _check(load().orc2d_sample(self.h, len(receivers), rec, len(types), tcodes,
                                   np.ascontiguousarray(fields, dtype=np.float64), out))
        return out

    def properties(self):
        """(rho, kappa, mu, w_i w_j J), each (nspec, 5, 5) in mesh element order."""
        out = [np.zeros((self.nspec, 5, 5)) for _ in range(4)]
        load().orc2d_properties(self.h, *out)
        return tuple(out)

    def set_start_time(self, t: float):
        """User start time for later runs (0: automatic, the earliest source onset)."""
        load().orc2d_set_start_time(self.h, float(t))

    def set_properties(self, rho, kappa, mu):
        """Replace the per-point properties ((nspec, 5, 5), mesh order) and rebuild the mass matrix."""
        c = [np.ascontiguousarray(x, dtype=np.float64) for x in (rho, kappa, mu)]
        load().orc2d_set_properties(self.h, *c)

    @property
    def nstacey(self) -> int:
        return load().orc2d_nstacey(self.h)

    def forward_record(self, sources, receivers, nstep: int, dt: float, nsb: int = 1,
                       types=("displacement",), nthreads: int = 1, snapshot_every: int = 0):
        """Forward run that keeps what a combined run needs: returns
        (seismograms, t0, final fields (3, 2, nglob), Stacey-element record
        (nstep, nstacey, 5, 5, 2)) and, with snapshot_every > 0, the
        checkpoints (nsnap, 3, 2, nglob) of steps 0, k, 2k, ... < nstep - 1 as a
        fifth item."""
        src, rec = self._src_rows(sources), self._rec_rows(receivers)
        tcodes = np.array([SEISMO_TYPES[t] for t in types], dtype=np.int32)
        nsig = -(-nstep // nsb)
        seis = np.zeros((nsig, len(types), len(receivers), 2))
        t0 = C.c_double()
        fields = np.zeros((3, 2, self.nglob))
        record = np.zeros((nstep, self.nstacey, 5, 5, 2))
        nsnap = (nstep - 2) // snapshot_every + 1 if snapshot_every > 0 else 0
        snaps = np.zeros((max(nsnap, 1), 3, 2, self.nglob))
        _check(load().orc2d_forward_record(self.h, len(sources), src, len(receivers), rec, nstep, dt, nsb,
                                           len(types), tcodes, seis, C.byref(t0), nthreads, fields, record,
                                           snapshot_every, snaps))
        if snapshot_every > 0:
            return seis, t0.value, fields, record, snaps[:nsnap]
        return seis, t0.value, fields, record

    def combined(self, sources, adjoint_positions, adjoint_stf, receivers, nstep: int, dt: float, fields, record,
                 nsb: int = 1, types=("displacement",), nthreads: int = 1, snapshot_every: int = 0,
                 snapshots=None):
        """Combined adjoint + backward run; adjoint_stf (nadj, nstep, 2) in
        forward time order.  Returns (kernels (6, nspec, 5, 5) mesh order: rho,
        mu, kappa, rhop, alpha, beta; backward seismograms, latest first)."""
        src, rec = self._src_rows(sources), self._rec_rows(receivers)
        adj = self._rec_rows(adjoint_positions)
        stf = np.ascontiguousarray(adjoint_stf, dtype=np.float64)
        assert stf.shape == (len(adjoint_positions), nstep, 2)
        tcodes = np.array([SEISMO_TYPES[t] for t in types], dtype=np.int32)
        nsig = -(-nstep // nsb)
        seis = np.zeros((nsig, len(types), len(receivers), 2))
        nspec = self.nspec
        kernels = np.zeros((6, nspec, 5, 5))
        _check(load().orc2d_combined(self.h, len(sources), src, len(adjoint_positions), adj, stf, len(receivers),
                                     rec, nstep, dt, nsb, len(types), tcodes,
                                     np.ascontiguousarray(fields, dtype=np.float64),
                                     np.ascontiguousarray(record, dtype=np.float64), kernels, seis, nthreads,
                                     snapshot_every,
                                     np.ascontiguousarray(snapshots if snapshots is not None else np.zeros(1),
                                                          dtype=np.float64)))
        return kernels, seis

    def forward(self, sources, receivers, nstep: int, dt: float, nsb: int = 1,
                types=("displacement",), nthreads: int = 1, return_fields: bool = False):
        """Run the forward loop; returns (seismograms (nsig, ntypes, nrec, 2), t0[, fields]).
        fields: (3, 2, nglob) u, v, a; on a mesh with acoustic elements the
        pair (elastic (3, 2, nglob_el), acoustic (3, nglob_ac) chi, chi_dot,
        chi_dot_dot)."""
        src, rec = self._src_rows(sources), self._rec_rows(receivers)
        tcodes = np.array([SEISMO_TYPES[t] for t in types], dtype=np.int32)
        nsig = nstep // nsb if nstep % nsb == 0 else nstep // nsb + 1
        seis = np.zeros((nsig, len(types), len(receivers), 2))
        t0 = C.c_double()
        coupled = self.nglob_ac > 0
        nf = 6 * self.nglob_el + 3 * self.nglob_ac if coupled else 6 * self.nglob
        fields = np.zeros(nf) if return_fields else None
        _check(load().orc2d_forward(self.h, len(sources), np.ascontiguousarray(src), len(receivers),
                                    np.ascontiguousarray(rec), nstep, dt, nsb, len(types), tcodes, seis,
                                    C.byref(t0), nthreads, _ptr(fields)))
        if return_fields:
            if coupled:
                ne = 6 * self.nglob_el
                fields = (fields[:ne].reshape(3, 2, self.nglob_el), fields[ne:].reshape(3, self.nglob_ac))
            else:
                fields = fields.reshape(3, 2, self.nglob)
            return seis, t0.value, fields
        return seis, t0.value

    def apply_stiffness(self, u, v):
        out = np.zeros((2, self.nglob))
        load().orc2d_apply_stiffness(self.h, np.ascontiguousarray(u, dtype=np.float64),
                                     np.ascontiguousarray(v, dtype=np.float64), out)
        return out


class Oracle3D:
    """3-D restatement on a structured hexahedral mesh (dict as built by
    ``specfempp_amd.meshes.hex_mesh``)."""

    def __init__(self, mesh, precision: int = 8, dt: float = 0.0):
        lib = load()
        self.precision = precision
        mat = np.ascontiguousarray(mesh["material"], dtype=np.int32).ravel()
        rho = np.ascontiguousarray(mesh["rho"], dtype=np.float64)
        vp = np.ascontiguousarray(mesh["vp"], dtype=np.float64)
        vs = np.ascontiguousarray(mesh["vs"], dtype=np.float64)
        absorbing = np.ascontiguousarray(mesh["absorbing"], dtype=np.int32)
        h = C.c_void_p()
        rc = lib.orc3d_create(precision, mesh["nx"], mesh["ny"], mesh["nz"], mesh.get("x0", 0.0), mesh.get("y0", 0.0),
                              mesh.get("z0", 0.0), mesh["hx"], mesh["hy"], mesh["hz"], mat, rho, vp, vs, absorbing,
                              dt, C.byref(h))
        if rc != 0:
            raise RuntimeError(lib.orc3d_last_error().decode())
        self.h = h
        self.nspec = mesh["nx"] * mesh["ny"] * mesh["nz"]
        self.nglob = lib.orc3d_nglob(h)

    @classmethod
    def from_database(cls, db, precision: int = 8, dt: float = 0.0):
        """3-D restatement on a mesh database (specfempp_amd.io.database3d
        layout): general hexahedra with per-point geometry."""
        lib = load()
        self = cls.__new__(cls)
        self.precision = precision
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32).reshape(-1)  # noqa: E731
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).reshape(-1)  # noqa: E731
        names = ("xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz")
        d9 = [f32(db.derivs[n]) for n in names]
        d9p = (C.c_void_p * 9)(*[a.ctypes.data for a in d9])
        ab = db.absorbing
        keep = dict(ibool=i32(db.ibool), x=f32(db.x), y=f32(db.y), z=f32(db.z), irr=i32(db.irregular_element_number),
                    jac=f32(db.derivs["jacobian"]), kappa=f32(db.kappa), mu=f32(db.mu), rho=f32(db.rho),
                    rvp=f32(db.rho_vp), rvs=f32(db.rho_vs), ai=i32(ab.ispec), aijk=i32(ab.ijk),
                    ajw=f32(ab.jacobian2Dw), an=f32(ab.normal))
        h = C.c_void_p()
        rc = lib.orc3d_create_db(precision, db.nspec, db.nglob, keep["ibool"], _ptr(keep["x"]), _ptr(keep["y"]),
                                 _ptr(keep["z"]), len(keep["irr"]), keep["irr"], float(db.xix_regular),
                                 float(db.jacobian_regular), d9p, _ptr(keep["jac"]), _ptr(keep["kappa"]),
                                 _ptr(keep["mu"]), _ptr(keep["rho"]), _ptr(keep["rvp"]), _ptr(keep["rvs"]),
                                 len(keep["ai"]), keep["ai"], keep["aijk"], _ptr(keep["ajw"]), _ptr(keep["an"]),
                                 dt, C.byref(h))
        if rc != 0:
            raise RuntimeError(lib.orc3d_last_error().decode())
        self.h = h
        self.nspec = db.nspec
        self.nglob = lib.orc3d_nglob(h)
        return self

    def __del__(self):
        if getattr(self, "h", None):
            load().orc3d_destroy(self.h)
            self.h = None

    def ibool(self):
        out = np.zeros(self.nspec * 125, dtype=np.int32)
        load().orc3d_ibool(self.h, out)
        return out.reshape(self.nspec, 5, 5, 5)

    def mass_inverse(self):
        out = np.zeros((3, self.nglob))
        load().orc3d_mass_inverse(self.h, out)
        return out

    def forward(self, sources, receivers, nstep, dt, nsb=1, types=("displacement",), return_fields=False,
                nthreads=1):
        """nthreads > 1 runs the element loop on that many OpenMP threads (same
        bits as the serial loop: every point still adds its element
        contributions in increasing element order)."""
        load().orc3d_set_threads(int(nthreads))
        src = np.zeros((len(sources), 16))
        for i, s in enumerate(sources):
            src[i] = [1 if s.get("type", "force") == "moment_tensor" else 0, s["x"], s.get("y", 0.0), s["z"],
                      s.get("fx", 0.0), s.get("fy", 0.0), s.get("fz", 0.0), s.get("mxx", 0.0), s.get("myy", 0.0),
                      s.get("mzz", 0.0), s.get("mxy", 0.0), s.get("mxz", 0.0), s.get("myz", 0.0), s["f0"],
                      s.get("tshift", 0.0), s.get("factor", 1.0)]
        rec = np.array([[r["x"], r.get("y", 0.0), r["z"]] for r in receivers], dtype=np.float64).reshape(-1, 3)
        tcodes = np.array([SEISMO_TYPES[t] for t in types], dtype=np.int32)
        nsig = (nstep + nsb - 1) // nsb
        seis = np.zeros((nsig, len(types), len(receivers), 3))
        t0 = C.c_double()
        fields = np.zeros((3, 3, self.nglob)) if return_fields else None
        rc = load().orc3d_forward(self.h, len(sources), np.ascontiguousarray(src), len(receivers),
                                  np.ascontiguousarray(rec), nstep, dt, nsb, len(types), tcodes, seis, C.byref(t0),
                                  _ptr(fields))
        if rc != 0:
            raise RuntimeError(load().orc3d_last_error().decode())
        if return_fields:
            return seis, t0.value, fields
        return seis, t0.value

    def apply_stiffness(self, u):
        out = np.zeros((3, self.nglob))
        load().orc3d_apply_stiffness(self.h, np.ascontiguousarray(u, dtype=np.float64), out)
        return out
