"""Reader and writer of the 3-D SPECFEM mesh database that the reference reads
(``specfem::io::read_3d_mesh``, src/io/mesh/impl/fortran/dim3/mesh.cpp:43-797):
a mesh parameters file (read_parameters.cpp:7-245) and a database file, both
sequences of Fortran sequential-unformatted records (a 4-byte length, the
payload, the length again; int, float and logical 4 bytes,
include/io/fortranio/fortran_io.tpp:44-58).

Record layout, as the reader's array helpers consume it
(include/io/mesh/impl/fortran/dim3/utilities.hpp):
- a scalar, or a 1-D array, is one record;
- an array of rank r > 1 is one record per leading index, its other indices
  Fortran-ordered inside the record (the second index fastest);
- index arrays are stored 1-based and read 0-based.

Arrays here follow the engine's element-point order p = (k * 5 + j) * 5 + i
(i along x fastest): per-element arrays are (nspec, 5, 5, 5) as [e, k, j, i],
which is the record order of the reference's (nspec, ngllx, nglly, ngllz)
views.  Face arrays keep the reference's shapes: ijk (nfaces, 3, 25), normal
(nfaces, 3, 25), jacobian2Dw (nfaces, 25).

``read_database_3d`` reads every section the reference reads (acoustic,
elastic and poroelastic arrays, coupling faces, MPI interfaces, anisotropy,
inner/outer elements, coloring, surface, adjacency), with the same
check values; ``write_database_3d`` writes the same layout.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

PARAM_ORDER = [
    (9999, ["acoustic_simulation", "elastic_simulation", "poroelastic_simulation", "anisotropy", "stacey_abc",
            "pml_abc", "approximate_ocean_load", "use_mesh_coloring"]),
    (9998, ["ndim", "ngllx", "nglly", "ngllz", "ngllsquare", "nproc"]),
    (9997, ["nspec", "nspec_poro", "nglob", "nglob_ocean"]),
    (9996, ["nspec2D_bottom", "nspec2D_top", "nspec2D_xmin", "nspec2D_xmax", "nspec2D_ymin", "nspec2D_ymax",
            "nspec_irregular"]),
    (9995, ["num_neighbors", "nfaces_surface", "num_abs_boundary_faces", "num_free_surface_faces",
            "num_coupling_ac_el_faces", "num_coupling_ac_po_faces", "num_coupling_el_po_faces",
            "num_coupling_po_el_faces", "num_interfaces_ext_mesh", "max_nibool_interfaces_ext_mesh"]),
    (9994, ["nspec_inner_acoustic", "nspec_outer_acoustic", "nspec_inner_elastic", "nspec_outer_elastic",
            "nspec_inner_poroelastic", "nspec_outer_poroelastic"]),
    (9993, ["num_phase_ispec_acoustic", "num_phase_ispec_elastic", "num_phase_ispec_poroelastic",
            "num_colors_inner_acoustic", "num_colors_outer_acoustic", "num_colors_inner_elastic",
            "num_colors_outer_elastic"]),
]
BOOL_PARAMS = set(PARAM_ORDER[0][1])
DERIVS = ("xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz")
ANISO = ("c11", "c12", "c13", "c14", "c15", "c16", "c22", "c23", "c24", "c25", "c26", "c33", "c34", "c35", "c36",
         "c44", "c45", "c46", "c55", "c56", "c66")
PORO = ("rho", "kappa", "eta", "tort", "perm", "phi", "rho_vpI", "rho_vpII", "rho_vsI")


@dataclass
class Faces:
    """Faces of a boundary or interface (ispec, ijk, jacobian2Dw, normal)."""
    ispec: np.ndarray
    ijk: np.ndarray          # (n, 3, 25) 0-based
    jacobian2Dw: np.ndarray  # (n, 25)
    normal: np.ndarray       # (n, 3, 25)

    @staticmethod
    def empty():
        return Faces(np.zeros(0, np.int32), np.zeros((0, 3, 25), np.int32), np.zeros((0, 25), np.float32),
                     np.zeros((0, 3, 25), np.float32))


@dataclass
class Mesh3DDatabase:
    params: dict
    ibool: np.ndarray                    # (nspec, 5, 5, 5) 0-based
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray                        # (nglob)
    irregular_element_number: np.ndarray  # (nspec_irregular) 0-based
    xix_regular: float
    jacobian_regular: float
    derivs: dict                         # name -> (nspec, 5, 5, 5), DERIVS and "jacobian"
    kappa: np.ndarray
    mu: np.ndarray
    rho: np.ndarray                      # (nspec, 5, 5, 5)
    ispec_type: np.ndarray               # (nspec) 0 acoustic, 1 elastic, 2 poroelastic
    mass_elastic: np.ndarray = None      # (nglob)
    mass_acoustic: np.ndarray = None
    rho_vp: np.ndarray = None            # (nspec, 5, 5, 5)
    rho_vs: np.ndarray = None
    poro: dict = field(default_factory=dict)
    absorbing: Faces = None
    abs_mass_elastic: tuple = None       # (x, y, z) each (nglob)
    abs_mass_acoustic: np.ndarray = None
    ibelm: dict = field(default_factory=dict)  # xmin ... top -> element lists
    free_surface: Faces = None
    coupling: dict = field(default_factory=dict)  # "ac_el", "ac_po", "el_po", "po_el" -> Faces
    mpi: dict = field(default_factory=dict)
    aniso: dict = field(default_factory=dict)
    ispec_is_inner: np.ndarray = None
    inner_outer: dict = field(default_factory=dict)  # medium -> (nspec_inner, nspec_outer, phase_ispec_inner)
    coloring: dict = field(default_factory=dict)     # medium -> (num_colors_outer, num_colors_inner, elements)
    nfaces_surface: int = 0               # as the surface section states it
    ispec_is_surface: np.ndarray = None
    iglob_is_surface: np.ndarray = None
    neighbors_xadj: np.ndarray = None
    neighbors_adjncy: np.ndarray = None

    @property
    def nspec(self):
        return int(self.params["nspec"])

    @property
    def nglob(self):
        return int(self.params["nglob"])


class _Reader:
    def __init__(self, data: bytes, name: str):
        self.d, self.o, self.name = data, 0, name

    def record(self) -> bytes:
        if self.o + 4 > len(self.d):
            raise ValueError(f"{self.name}: unexpected end of file")
        (n,) = struct.unpack_from("<i", self.d, self.o)
        payload = self.d[self.o + 4:self.o + 4 + n]
        (m,) = struct.unpack_from("<i", self.d, self.o + 4 + n)
        if m != n or len(payload) != n:
            raise ValueError(f"{self.name}: corrupt record at byte {self.o}")
        self.o += n + 8
        return payload

    def values(self, dtype, count=None):
        rec = self.record()
        arr = np.frombuffer(rec, dtype=dtype)
        if count is not None and arr.size != count:
            raise ValueError(f"{self.name}: record of {arr.size} values, expected {count}")
        return arr

    def scalar(self, dtype="<i4"):
        return self.values(dtype, 1)[0]

    def scalars(self, n, dtype="<i4"):
        return self.values(dtype, n)

    def check(self, value):
        v = int(self.scalar())
        if v != value:
            raise ValueError(f"{self.name}: test value ({value}) != read value ({v})")

    def array(self, shape, dtype="<f4", index=False):
        """A rank-r array: one record (r = 1) or one record per leading index
        holding the other indices Fortran-ordered."""
        shape = tuple(int(s) for s in shape)
        if len(shape) == 1:
            out = self.values(dtype, shape[0]).copy()
        else:
            inner = int(np.prod(shape[1:]))
            out = np.empty(shape, dtype=np.dtype(dtype).newbyteorder("="))
            for i in range(shape[0]):
                out[i] = self.values(dtype, inner).reshape(shape[1:][::-1]).T
        return out - 1 if index else out

    def elem(self, nspec, dtype="<f4", index=False):
        """(nspec, 5, 5, 5) as [e, k, j, i]: one record per element, x fastest."""
        out = np.empty((nspec, 5, 5, 5), dtype=np.dtype(dtype).newbyteorder("="))
        for e in range(nspec):
            out[e] = self.values(dtype, 125).reshape(5, 5, 5)
        return out - 1 if index else out


class _Writer:
    def __init__(self):
        self.parts = []

    def record(self, payload: bytes):
        self.parts += [struct.pack("<i", len(payload)), payload, struct.pack("<i", len(payload))]

    def values(self, arr, dtype):
        self.record(np.ascontiguousarray(arr, dtype=dtype).tobytes())

    def scalar(self, v, dtype="<i4"):
        self.values(np.array([v]), dtype)

    def array(self, arr, dtype="<f4", index=False):
        a = np.asarray(arr)
        if index:
            a = a.astype(np.int64) + 1
        if a.ndim == 1:
            self.values(a, dtype)
        else:
            for i in range(a.shape[0]):
                self.values(a[i].T.ravel(), dtype)

    def elem(self, arr, dtype="<f4", index=False):
        a = np.asarray(arr)
        if index:
            a = a.astype(np.int64) + 1
        for e in range(a.shape[0]):
            self.values(a[e].ravel(), dtype)

    def bytes(self):
        return b"".join(self.parts)


def _read_faces(r: _Reader, n: int) -> Faces:
    return Faces(r.array((n,), "<i4", index=True), r.array((n, 3, 25), "<i4", index=True),
                 r.array((n, 25)), r.array((n, 3, 25)))


def _write_faces(w: _Writer, f: Faces):
    w.array(f.ispec, "<i4", index=True)
    w.array(f.ijk, "<i4", index=True)
    w.array(f.jacobian2Dw)
    w.array(f.normal)


def read_parameters_3d(path: str) -> dict:
    """The mesh parameters file (read_parameters.cpp:7-245)."""
    r = _Reader(open(path, "rb").read(), path)
    params = {}
    for check, names in PARAM_ORDER:
        r.check(check)
        for n in names:
            v = int(r.scalar())
            params[n] = bool(v & 0xFF) if n in BOOL_PARAMS else v
    r.check(9992)
    return params


def write_parameters_3d(path: str, params: dict):
    w = _Writer()
    for check, names in PARAM_ORDER:
        w.scalar(check)
        for n in names:
            w.scalar(int(bool(params[n])) if n in BOOL_PARAMS else int(params[n]))
    w.scalar(9992)
    open(path, "wb").write(w.bytes())


def read_database_3d(parameters_path: str, database_path: str) -> Mesh3DDatabase:
    """Both files, in the order of read_3d_mesh (mesh.cpp:43-797)."""
    P = read_parameters_3d(parameters_path)
    r = _Reader(open(database_path, "rb").read(), database_path)

    def check_value(name, v):
        if int(v) != int(P[name]):
            raise ValueError(f"{name} value ({int(v)}) != expected value ({int(P[name])})")

    nspec, nglob, nirr = int(r.scalar()), int(r.scalar()), int(r.scalar())
    check_value("nspec", nspec)
    check_value("nglob", nglob)
    check_value("nspec_irregular", nirr)
    ibool = r.elem(nspec, "<i4", index=True)
    x, y, z = r.array((nglob,)), r.array((nglob,)), r.array((nglob,))
    irr = r.array((nirr,), "<i4", index=True)
    xix_regular = float(r.scalar("<f4"))
    jacobian_regular = float(r.scalar("<f4"))
    derivs = {n: r.elem(nspec) for n in DERIVS + ("jacobian",)}
    r.check(10000)
    kappa, mu = r.elem(nspec), r.elem(nspec)
    nac, nel, npo = int(r.scalar()), int(r.scalar()), int(r.scalar())
    ispec_type = r.array((nspec,), "<i4")
    if not (np.sum(ispec_type == 0) == nac and np.sum(ispec_type == 1) == nel and np.sum(ispec_type == 2) == npo):
        raise ValueError("element type counts do not match ispec_type")
    if np.any((ispec_type < 0) | (ispec_type > 2)):
        raise ValueError("Spectral element is not assigned to any material")
    r.check(9999)
    db = Mesh3DDatabase(P, ibool, x, y, z, irr, xix_regular, jacobian_regular, derivs, kappa, mu, None, ispec_type)
    if P["acoustic_simulation"]:
        db.mass_acoustic = r.array((nglob,))
    db.rho = r.elem(nspec)
    r.check(9998)
    if P["elastic_simulation"]:
        db.mass_elastic = r.array((nglob,))
        db.rho_vp = r.elem(nspec)
        db.rho_vs = r.elem(nspec)
    r.check(9997)
    if P["poroelastic_simulation"]:
        db.poro["mass_solid"] = r.array((nglob,))
        db.poro["mass_fluid"] = r.array((nglob,))
        inner = {"rho": 2, "kappa": 3, "perm": 6}  # materials.hpp:160-170
        for n in PORO:
            db.poro[n] = r.array((nspec, inner[n], 5, 5, 5)) if n in inner else r.elem(nspec)
    r.check(9996)
    nabs = int(r.scalar())
    check_value("num_abs_boundary_faces", nabs)
    if nabs > 0:
        db.absorbing = _read_faces(r, nabs)
        if P["elastic_simulation"]:
            db.abs_mass_elastic = tuple(r.array((nglob,)) for _ in range(3))
        if P["acoustic_simulation"]:
            db.abs_mass_acoustic = r.array((nglob,))
    else:
        db.absorbing = Faces.empty()
    r.check(9995)
    sides = ("xmin", "xmax", "ymin", "ymax", "bottom", "top")
    counts = [int(r.scalar()) for _ in sides]
    for s, c in zip(sides, counts):
        check_value(f"nspec2D_{s}", c)
    if P["num_abs_boundary_faces"] > 0:
        for s, c in zip(sides, counts):
            db.ibelm[s] = r.array((c,), "<i4", index=True) if c > 0 else np.zeros(0, np.int32)
    nfs = int(r.scalar())
    check_value("num_free_surface_faces", nfs)
    db.free_surface = _read_faces(r, nfs)
    for key, pname in (("ac_el", "num_coupling_ac_el_faces"), ("ac_po", "num_coupling_ac_po_faces"),
                       ("el_po", "num_coupling_el_po_faces")):
        n = int(r.scalar())
        check_value(pname, n)
        if n > 0:
            db.coupling[key] = _read_faces(r, n)
            if key == "el_po":
                db.coupling["po_el"] = _read_faces(r, int(P["num_coupling_po_el_faces"]))
    r.check(9997)
    nint, maxnib = int(r.scalar()), int(r.scalar())
    check_value("num_interfaces_ext_mesh", nint)
    check_value("max_nibool_interfaces_ext_mesh", maxnib)
    if nint > 0:
        db.mpi["neighbors"] = r.array((nint,), "<i4", index=True)
        db.mpi["nibool"] = r.array((nint,), "<i4", index=True)
        db.mpi["ibool"] = r.array((nint, maxnib), "<i4", index=True)
    r.check(9996)
    if P["anisotropy"]:
        for n in ANISO:
            db.aniso[n] = r.elem(nspec)
    r.check(9995)
    db.ispec_is_inner = r.array((nspec,), "<i4").astype(bool)
    for medium, flag, check in (("acoustic", "acoustic_simulation", 9994), ("elastic", "elastic_simulation", 9993),
                                ("poroelastic", "poroelastic_simulation", 9992)):
        if P[flag]:
            ni, no = (int(v) for v in r.scalars(2))
            nph = int(r.scalar())
            phase = r.array((nph, 2), "<i4", index=True) if nph > 0 else np.zeros((0, 2), np.int32)
            db.inner_outer[medium] = (ni, no, phase)
        r.check(check)
    if P["use_mesh_coloring"]:
        # both blocks are guarded by the acoustic flag in the reference (mesh.cpp:718-750)
        for medium in ("acoustic", "elastic"):
            if P["acoustic_simulation"]:
                nco, nci = int(r.scalar()), int(r.scalar())
                db.coloring[medium] = (nco, nci, r.array((nco + nci,), "<i4", index=True))
    r.check(9991)
    nsurf = int(r.scalar())
    db.nfaces_surface = nsurf
    db.ispec_is_surface = r.array((nspec,), "<i4").astype(bool)
    db.iglob_is_surface = r.array((nglob,), "<i4").astype(bool)
    r.check(9990)
    nnb = int(r.scalar())
    db.neighbors_xadj = r.array((nspec + 1,), "<i4", index=True)
    db.neighbors_adjncy = r.array((nnb,), "<i4", index=True)
    r.check(9989)
    if r.o != len(r.d):
        raise ValueError(f"{database_path}: {len(r.d) - r.o} bytes after the last record")
    return db


def write_database_3d(parameters_path: str, database_path: str, db: Mesh3DDatabase):
    """Write both files in the layout read_database_3d (and the reference) reads."""
    P = db.params
    write_parameters_3d(parameters_path, P)
    w = _Writer()
    nspec, nglob = db.nspec, db.nglob
    w.scalar(nspec)
    w.scalar(nglob)
    w.scalar(len(db.irregular_element_number))
    w.elem(db.ibool, "<i4", index=True)
    for a in (db.x, db.y, db.z):
        w.array(a)
    w.array(db.irregular_element_number, "<i4", index=True)
    w.scalar(db.xix_regular, "<f4")
    w.scalar(db.jacobian_regular, "<f4")
    for n in DERIVS + ("jacobian",):
        w.elem(db.derivs[n])
    w.scalar(10000)
    w.elem(db.kappa)
    w.elem(db.mu)
    for code in (0, 1, 2):
        w.scalar(int(np.sum(db.ispec_type == code)))
    w.array(db.ispec_type, "<i4")
    w.scalar(9999)
    if P["acoustic_simulation"]:
        w.array(db.mass_acoustic)
    w.elem(db.rho)
    w.scalar(9998)
    if P["elastic_simulation"]:
        w.array(db.mass_elastic)
        w.elem(db.rho_vp)
        w.elem(db.rho_vs)
    w.scalar(9997)
    if P["poroelastic_simulation"]:
        raise ValueError("writing poroelastic databases is not implemented")
    w.scalar(9996)
    nabs = len(db.absorbing.ispec)
    w.scalar(nabs)
    if nabs > 0:
        _write_faces(w, db.absorbing)
        if P["elastic_simulation"]:
            for a in db.abs_mass_elastic:
                w.array(a)
        if P["acoustic_simulation"]:
            w.array(db.abs_mass_acoustic)
    w.scalar(9995)
    sides = ("xmin", "xmax", "ymin", "ymax", "bottom", "top")
    for s in sides:
        w.scalar(len(db.ibelm.get(s, ())))
    if nabs > 0:
        for s in sides:
            if len(db.ibelm.get(s, ())) > 0:
                w.array(db.ibelm[s], "<i4", index=True)
    w.scalar(len(db.free_surface.ispec))
    _write_faces(w, db.free_surface)
    for key in ("ac_el", "ac_po", "el_po"):
        f = db.coupling.get(key)
        w.scalar(0 if f is None else len(f.ispec))
        if f is not None and len(f.ispec) > 0:
            _write_faces(w, f)
            if key == "el_po":
                _write_faces(w, db.coupling["po_el"])
    w.scalar(9997)
    nint = int(P["num_interfaces_ext_mesh"])
    w.scalar(nint)
    w.scalar(int(P["max_nibool_interfaces_ext_mesh"]))
    if nint > 0:
        w.array(db.mpi["neighbors"], "<i4", index=True)
        w.array(db.mpi["nibool"], "<i4", index=True)
        w.array(db.mpi["ibool"], "<i4", index=True)
    w.scalar(9996)
    if P["anisotropy"]:
        for n in ANISO:
            w.elem(db.aniso[n])
    w.scalar(9995)
    w.array(np.asarray(db.ispec_is_inner, dtype=np.int32), "<i4")
    for medium, flag, check in (("acoustic", "acoustic_simulation", 9994), ("elastic", "elastic_simulation", 9993),
                                ("poroelastic", "poroelastic_simulation", 9992)):
        if P[flag]:
            ni, no, phase = db.inner_outer[medium]
            w.values(np.array([ni, no]), "<i4")
            w.scalar(len(phase))
            if len(phase) > 0:
                w.array(phase, "<i4", index=True)
        w.scalar(check)
    if P["use_mesh_coloring"]:
        for medium in ("acoustic", "elastic"):
            if P["acoustic_simulation"]:
                nco, nci, elems = db.coloring[medium]
                w.scalar(nco)
                w.scalar(nci)
                w.array(elems, "<i4", index=True)
    w.scalar(9991)
    w.scalar(int(db.nfaces_surface or P["nfaces_surface"]))
    w.array(np.asarray(db.ispec_is_surface, dtype=np.int32), "<i4")
    w.array(np.asarray(db.iglob_is_surface, dtype=np.int32), "<i4")
    w.scalar(9990)
    w.scalar(len(db.neighbors_adjncy))
    w.array(db.neighbors_xadj, "<i4", index=True)
    w.array(db.neighbors_adjncy, "<i4", index=True)
    w.scalar(9989)
    open(database_path, "wb").write(w.bytes())
