"""Reader for the 2-D SPECFEM mesh database (``database.bin``).

The file is a sequence of Fortran sequential-unformatted records: a 4-byte
little-endian length, the payload, and the same length again
(``include/io/fortranio/fortran_io.tpp:44-58``; record widths from
``setup/specfem_setup.hpp.in:7``: int 4, float 4, double 8, logical 4,
character 512).

The record order follows the reference reader:
  header ............ ``src/io/mesh/impl/fortran/dim2/read_mesh_database.cpp:13-157``
  control nodes ..... ``read_mesh_database.cpp:159-182``
  mesh parameters ... ``src/io/mesh/impl/fortran/dim2/read_parameters.cpp:5-59``
  attenuation ....... ``read_mesh_database.cpp:184-205``
  materials ......... ``src/io/mesh/impl/fortran/dim2/read_material_properties.cpp:160-232``
  element materials . ``read_material_properties.cpp:458-497``
  interfaces count .. ``src/io/mesh/impl/fortran/dim2/mesh.cpp:85``
  absorbing edges ... ``src/io/mesh/impl/fortran/dim2/read_boundaries.cpp:133-204``
  acoustic surface .. ``read_boundaries.cpp:206-268`` (element, edge or node)
  coupled interfaces  ``src/io/mesh/impl/fortran/dim2/read_interfaces.cpp:9-30``
                      (fluid-solid pairs: acoustic element, then elastic)

Only what the elastic P-SV and acoustic isotropic paths need is interpreted;
the rest is skipped record by record so that a malformed file still fails
loudly.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

# Edge codes of an absorbing edge, in the order the database lists the flags
# (codeabsread1..4 -> bottom, right, top, left;
# ``read_boundaries.cpp:166-179``).
EDGE_BOTTOM, EDGE_RIGHT, EDGE_TOP, EDGE_LEFT = 0, 1, 2, 3
# Corner codes of an acoustic free-surface entry that names one control node
# (``read_boundaries.cpp:218-232``).
CORNER_BOTTOM_LEFT, CORNER_BOTTOM_RIGHT, CORNER_TOP_RIGHT, CORNER_TOP_LEFT = 4, 5, 6, 7


@dataclass
class Material:
    """One material line (``read_material_properties.cpp:172-238``)."""
    indic: int          # 1 isotropic, 2 anisotropic, ...
    density: float
    cp: float
    cs: float
    qkappa: float
    qmu: float
    raw: tuple = ()

    @property
    def is_acoustic(self) -> bool:
        return self.indic == 1 and self.cs == 0.0 and self.raw[2] == 0.0


@dataclass
class Mesh2D:
    """Mesh-order arrays of a 2-D database (0-based indices)."""
    nspec: int
    npgeo: int
    ngnod: int
    coorg: np.ndarray            # (2, npgeo) float64 as stored in the file
    knods: np.ndarray            # (ngnod, nspec) int32
    kmato: np.ndarray            # (nspec,) int32 material index (0-based)
    materials: list = field(default_factory=list)
    abs_ispec: np.ndarray = None  # (nelemabs,) int32 mesh element of each edge
    abs_type: np.ndarray = None   # (nelemabs,) int32 EDGE_* code
    afs_ispec: np.ndarray = None  # (n,) int32 mesh element of each acoustic free-surface entry
    afs_type: np.ndarray = None   # (n,) int32 EDGE_* or CORNER_* code
    fs_acoustic: np.ndarray = None  # (n,) int32 acoustic element of each fluid-solid edge
    fs_elastic: np.ndarray = None   # (n,) int32 elastic element of the same edge
    nstep: int = 0
    dt: float = 0.0


class _Records:
    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0

    def next(self) -> bytes:
        if self.pos + 4 > len(self.data):
            raise ValueError("database.bin: unexpected end of file")
        (n,) = struct.unpack_from("<i", self.data, self.pos)
        start = self.pos + 4
        end = start + n
        if n < 0 or end + 4 > len(self.data):
            raise ValueError(f"database.bin: corrupt record length {n} at {self.pos}")
        (n2,) = struct.unpack_from("<i", self.data, end)
        if n2 != n:
            raise ValueError(f"database.bin: record markers differ ({n} vs {n2}) at {self.pos}")
        self.pos = end + 4
        return self.data[start:end]

    def peek_len(self) -> int:
        (n,) = struct.unpack_from("<i", self.data, self.pos)
        return n

    def read(self, fmt: str):
        rec = self.next()
        size = struct.calcsize("<" + fmt)
        if size != len(rec):
            raise ValueError(f"database.bin: record of {len(rec)} bytes, expected {size} ({fmt})")
        return struct.unpack("<" + fmt, rec)


# Header record formats in file order (read_mesh_database.cpp:24-152).  's' is
# a 512-byte character record.
_HEADER = [
    "512s", "ii", "i", "ii", "ii", "i", "i", "i", "i", "i", "d", "d", "d", "d",
    "d", "i", "i", "i", "iii", "ii", "iiidid", "d", "i", "iiii", "512si",
    "512s", "512s", "512s", "iiii", "i", "iii", "i", "i", "ii", "i", "i", "i",
    "i", "i", "i", "i", "dd",
]
# After the optional electromagnetic record (i, i, d) come these
# (read_mesh_database.cpp:124-152).
_HEADER_TAIL = [
    "i", "i", "d", "i", "d", "i", "i", "i", "d", "i", "d", "i", "i", "d", "i",
    "i", "id", "i", "i", "i", "i", "ii",
]


def read_database_2d(path: str) -> Mesh2D:
    """Parse a 2-D ``database.bin`` (``src/io/mesh/impl/fortran/dim2/mesh.cpp:26-193``);
    a path ending in ``.gz`` is read through gzip."""
    if str(path).endswith(".gz"):
        import gzip

        with gzip.open(path, "rb") as fh:
            r = _Records(fh.read())
    else:
        with open(path, "rb") as fh:
            r = _Records(fh.read())

    vals = []
    for fmt in _HEADER:
        vals.append(r.read(fmt))
    nspec = vals[2][0]
    npgeo, _nproc = vals[3]
    # The optional (ATTENUATION_PERMITTIVITY, ATTENUATION_CONDUCTIVITY,
    # f0_electromagnetic) record: the reference tries to read it and seeks back
    # when the record does not have exactly that size.
    if r.peek_len() == 16:
        r.read("iid")
    tail = [r.read(fmt) for fmt in _HEADER_TAIL]
    nstep, dt = tail[16]

    coorg = np.zeros((2, npgeo), dtype=np.float64)
    for _ in range(npgeo):
        ipoin, x, z = r.read("idd")
        if ipoin < 1 or ipoin > npgeo:
            raise ValueError("database.bin: control node index out of range")
        coorg[0, ipoin - 1] = x
        coorg[1, ipoin - 1] = z

    numat, ngnod, nspec_p, _pointsdisp, _plot = r.read("iiiii")
    if nspec_p != nspec:
        raise ValueError("database.bin: nspec mismatch between header and parameters")
    (nelemabs, nelem_acforcing, nelem_acoustic_surface, n_fs, n_fp, n_sp,
     nnodes_tangential, nelem_axis) = r.read("iiiiiiii")
    r.read("idi")  # N_SLS, attenuation f0, read_velocities_at_f0

    materials = []
    for i in range(numat):
        v = r.read("ii" + "d" * 13)
        n, indic = v[0], v[1]
        if n != i + 1:
            raise ValueError("database.bin: wrong material set number")
        d = v[2:]
        # isotropic: density, cp, cs, compaction, -, Qkappa, Qmu
        materials.append(Material(indic=indic, density=d[0], cp=d[1], cs=d[2],
                                  qkappa=d[5], qmu=d[6], raw=d))

    knods = np.zeros((ngnod, nspec), dtype=np.int32)
    kmato = np.zeros(nspec, dtype=np.int32)
    for _ in range(nspec):
        v = r.read("ii" + "i" * ngnod + "i")
        n, mat = v[0], v[1]
        if n < 1 or n > nspec or mat < 1 or mat > numat:
            raise ValueError("database.bin: element material record out of range")
        nodes = np.asarray(v[2:2 + ngnod], dtype=np.int32)
        if np.any(nodes == 0):
            raise ValueError("database.bin: zero control node index")
        knods[:, n - 1] = nodes - 1
        kmato[n - 1] = mat - 1

    r.read("ii")  # ninterfaces, max_interface_size

    abs_ispec = np.zeros(max(nelemabs, 0), dtype=np.int32)
    abs_type = np.zeros(max(nelemabs, 0), dtype=np.int32)
    for i in range(max(nelemabs, 0)):
        v = r.read("i" + "iiii" + "i" + "i" * 8)
        num, codes = v[0], v[1:5]
        if num < 1 or num > nspec:
            raise ValueError("database.bin: wrong absorbing element number")
        if sum(1 for c in codes if c) != 1:
            raise ValueError("database.bin: one and only one absorbing edge per line")
        abs_ispec[i] = num - 1
        abs_type[i] = [j for j, c in enumerate(codes) if c][0]

    afs_ispec = np.zeros(max(nelem_acoustic_surface, 0), dtype=np.int32)
    afs_type = np.zeros(max(nelem_acoustic_surface, 0), dtype=np.int32)
    for i in range(max(nelem_acoustic_surface, 0)):
        e, kind, n1, n2 = r.read("iiii")
        if e < 1 or e > nspec:
            raise ValueError("database.bin: acoustic free surface element out of range")
        afs_ispec[i] = e - 1
        afs_type[i] = _surface_type(kind, n1 - 1, n2 - 1, knods[:, e - 1])
    for _ in range(nelem_acforcing):
        r.next()
    fs_acoustic = np.zeros(max(n_fs, 0), dtype=np.int32)
    fs_elastic = np.zeros(max(n_fs, 0), dtype=np.int32)
    for i in range(max(n_fs, 0)):
        ea, ee = r.read("ii")
        if not (1 <= ea <= nspec and 1 <= ee <= nspec):
            raise ValueError("database.bin: fluid-solid interface element out of range")
        fs_acoustic[i] = ea - 1
        fs_elastic[i] = ee - 1
    for _ in range(n_fp + n_sp):
        r.next()
    r.next()  # tangential flags
    for _ in range(max(nnodes_tangential, 0)):
        r.next()
    for _ in range(nelem_axis):
        r.next()

    return Mesh2D(nspec=nspec, npgeo=npgeo, ngnod=ngnod, coorg=coorg,
                  knods=knods, kmato=kmato, materials=materials,
                  abs_ispec=abs_ispec, abs_type=abs_type, afs_ispec=afs_ispec, afs_type=afs_type,
                  fs_acoustic=fs_acoustic, fs_elastic=fs_elastic, nstep=nstep, dt=dt)


def _surface_type(kind: int, n1: int, n2: int, nodes) -> int:
    """Edge or corner of an acoustic free-surface entry from its control nodes
    (0-based), ``read_boundaries.cpp:206-248``: kind 1 names one node (a
    corner), otherwise two (an edge)."""
    c = [int(x) for x in nodes[:4]]
    if kind == 1:
        for code, node in zip((CORNER_BOTTOM_LEFT, CORNER_BOTTOM_RIGHT, CORNER_TOP_RIGHT, CORNER_TOP_LEFT), c):
            if n1 == node:
                return code
    else:
        pairs = ((EDGE_BOTTOM, c[0], c[1]), (EDGE_LEFT, c[0], c[3]), (EDGE_RIGHT, c[1], c[2]), (EDGE_TOP, c[2], c[3]))
        for code, a, b in pairs:
            if {n1, n2} == {a, b}:
                return code
    raise ValueError("database.bin: could not classify an acoustic free surface entry")


def read_fortran_array(path: str, dtype) -> np.ndarray:
    """Read a file of one-value Fortran records (the reference's test arrays,
    ``tests/unit-tests/utilities/include/compare_array.tpp:139-158``)."""
    dt = np.dtype(dtype)
    raw = np.fromfile(path, dtype=np.uint8)
    rec = 8 + dt.itemsize
    if raw.size % rec:
        raise ValueError(f"{path}: size {raw.size} is not a multiple of {rec}")
    n = raw.size // rec
    view = raw.reshape(n, rec)
    lens = view[:, :4].copy().view("<i4").ravel()
    if np.any(lens != dt.itemsize):
        raise ValueError(f"{path}: unexpected record length")
    return view[:, 4:4 + dt.itemsize].copy().view(dt).ravel()


def read_stations(path: str):
    """STATIONS file: name, network, x, z, ... (``src/io/receivers.cpp:22-50``)."""
    out = []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) < 4:
                continue
            out.append((parts[0], parts[1], float(parts[2]), float(parts[3])))
    return out


def read_trace(path: str) -> np.ndarray:
    """ASCII seismogram: two columns (time, value) per line."""
    return np.loadtxt(path, dtype=np.float64)
