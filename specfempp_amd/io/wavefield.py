"""Wavefield, boundary-value and kernel files in the reference's ASCII layout.

The reference's ASCII I/O library (include/io/ASCII/impl/*) maps a file to a
directory, a group to a subdirectory and a dataset to ``<name>.txt`` (one
value per line, floats as ``std::scientific`` with 10 digits, integers plain,
in the memory order of the Kokkos view) beside ``<name>.meta`` (``rank R`` then
``dims d0 d1 ...``).  On top of it:

* ForwardWavefield (include/io/wavefield/writer.tpp:8-148, reader.tpp:13-63):
  ``Coordinates/<medium>/{X, Z, mapping}`` at the start of a forward run, and
  ``Step%06d/<medium>/{Displacement, Velocity, Acceleration}`` (acoustic:
  ``Potential, PotentialDot, PotentialDotDot``) plus
  ``Step%06d/Boundary/Stacey/{IndexMapping, ElasticAcceleration,
  AcousticAcceleration, PoroelasticAcceleration}`` when the writer task runs
  (by default once, after the last step, as ``Step<nstep>``).  The combined
  run reads the step back into its buffer wavefield and boundary values.
* Kernels (include/io/kernel/writer.tpp:14, io/impl/medium_writer.tpp:12-70):
  ``Kernels/<medium>_<property>/{X, Z, rho, mu, kappa, rhop, alpha, beta}``
  over the elements of that medium in compute order, each (n, ngllz, ngllx)
  in the DomainView layout of a one-element chunk (iz fastest, then ix).

Views in LayoutLeft are written column-major (``order="F"``).  The HDF5
backend is not available here (no h5py); files are the ASCII backend's.
Only media present in the mesh get groups with data: this engine's 2-D
meshes are elastic P-SV, so the other media's field groups are written empty
(as the reference writes them for a mesh without those media) and their
kernel groups are left out.
"""
from __future__ import annotations

import os
import shutil

import numpy as np

# medium tags and their field components (specfem::element::attributes)
COMPONENTS = {"elastic_psv": 2, "elastic_sh": 1, "acoustic": 1, "poroelastic": 4, "elastic_psv_t": 3}
STEP_MEDIA = ("elastic_psv", "elastic_sh", "acoustic", "poroelastic", "elastic_psv_t")
COORD_MEDIA = ("elastic_psv", "elastic_psv_t", "elastic_sh", "acoustic", "poroelastic")
KERNEL_NAMES = ("rho", "mu", "kappa", "rhop", "alpha", "beta")


# ---- the ASCII library ---------------------------------------------------------
def fresh_dir(path: str) -> str:
    """A file or group: the directory is replaced if it exists (File<write>,
    Group<write> in include/io/ASCII/impl/{file,group}.hpp)."""
    if os.path.exists(path):
        shutil.rmtree(path)
    os.makedirs(path)
    return path


def write_dataset(folder: str, name: str, data: np.ndarray, order: str = "F") -> None:
    """``name.meta`` and ``name.txt`` of one dataset; ``order`` is the memory
    order of the view ("F" for LayoutLeft, "C" for LayoutRight)."""
    data = np.asarray(data)
    with open(os.path.join(folder, name + ".meta"), "w") as fh:
        fh.write(f"rank {data.ndim}\n")
        fh.write("dims " + "".join(f"{d} " for d in data.shape) + "\n")
    flat = data.ravel(order=order)
    with open(os.path.join(folder, name + ".txt"), "w") as fh:
        if np.issubdtype(flat.dtype, np.integer):
            fh.write("".join(f"{int(v)}\n" for v in flat))
        else:
            fh.write("".join(f"{float(v):.10e}\n" for v in flat))


def read_meta(folder: str, name: str):
    with open(os.path.join(folder, name + ".meta")) as fh:
        head = fh.readline().split()
        dims = fh.readline().split()
    if len(head) != 2 or head[0] != "rank" or not dims or dims[0] != "dims":
        raise ValueError(f"{folder}/{name}.meta is corrupted")
    shape = tuple(int(d) for d in dims[1:])
    if len(shape) != int(head[1]):
        raise ValueError(f"{folder}/{name}.meta: rank {head[1]} but {len(shape)} dims")
    return shape


def read_dataset(folder: str, name: str, dtype=np.float64, order: str = "F", shape=None) -> np.ndarray:
    """A dataset back as an array of its meta shape; ``shape`` (if given) must
    match it, as the reference's reader checks rank and dims
    (include/io/ASCII/impl/datasetbase.hpp)."""
    meta = read_meta(folder, name)
    if shape is not None and tuple(shape) != meta:
        raise ValueError(f"{folder}/{name}: dims {meta} do not match the expected {tuple(shape)}")
    n = int(np.prod(meta)) if meta else 1
    path = os.path.join(folder, name + ".txt")
    vals = np.loadtxt(path, dtype=dtype, ndmin=1) if n else np.zeros(0, dtype=dtype)
    if vals.size != n:
        raise ValueError(f"{path}: {vals.size} values, dims say {n}")
    return vals.reshape(meta, order=order)


# ---- engine-side quantities --------------------------------------------------
def _compute_order_fields(engine):
    """(coordinates (nspec, 5, 5, 2) compute order, index mapping, field index)."""
    return engine.point_coordinates(), engine.index_mapping(), engine.field_index()


def stacey_index_mapping(engine) -> np.ndarray:
    """boundary_values.stacey property_index_mapping (nspec, compute order):
    the Stacey-element index of each element, -1 elsewhere
    (boundary_medium_container.tpp:20-31; Stacey elements come last in
    compute order)."""
    nstacey = engine.stacey_elements()
    out = np.full(engine.nspec, -1, dtype=np.int32)
    out[engine.nspec - nstacey:] = np.arange(nstacey, dtype=np.int32)
    return out


# ---- ForwardWavefield -------------------------------------------------------
class WavefieldWriter:
    """Periodic task of a forward run writing ``<folder>/ForwardWavefield``
    (periodic_tasks/wavefield_writer.hpp): coordinates at initialize, the
    state and the Stacey record at every ``time_interval`` steps and, with
    ``include_last_step``, after the last step.  The engine must run with
    ``forward_record`` semantics (the record is what the file stores)."""

    def __init__(self, folder: str, time_interval: int = 0, include_last_step: bool = True):
        self.root = os.path.join(folder, "ForwardWavefield")
        self.time_interval, self.include_last_step = int(time_interval), bool(include_last_step)

    def should_run(self, istep: int) -> bool:
        if self.include_last_step and istep == -1:
            return True
        return self.time_interval > 0 and istep % self.time_interval == 0

    def initialize(self, engine):
        fresh_dir(self.root)
        base = fresh_dir(os.path.join(self.root, "Coordinates"))
        coords, im, fi = _compute_order_fields(engine)
        for medium in COORD_MEDIA:
            g = fresh_dir(os.path.join(base, medium))
            if medium == "elastic_psv":
                nglob = engine.nglob
                mapping = fi[im]  # (nspec, 5, 5) per-medium index of every point
                x = np.zeros(nglob, dtype=np.float32)
                z = np.zeros(nglob, dtype=np.float32)
                x[mapping.ravel()] = coords[..., 0].ravel()
                z[mapping.ravel()] = coords[..., 1].ravel()
            else:
                mapping, x, z = np.zeros((0, 5, 5), dtype=np.int32), np.zeros(0), np.zeros(0)
            write_dataset(g, "X", x)
            write_dataset(g, "Z", z)
            write_dataset(g, "mapping", mapping.astype(np.int32))

    def run(self, engine, istep: int):
        base = fresh_dir(os.path.join(self.root, f"Step{istep:06d}"))
        for medium in STEP_MEDIA:
            g = fresh_dir(os.path.join(base, medium))
            ncomp = COMPONENTS[medium]
            names = (("Potential", "PotentialDot", "PotentialDotDot") if medium == "acoustic"
                     else ("Displacement", "Velocity", "Acceleration"))
            for name, which in zip(names, ("displacement", "velocity", "acceleration")):
                if medium == "elastic_psv":
                    data = engine.wavefield_field("forward", which).T  # (nglob, ncomp) LayoutLeft
                else:
                    data = np.zeros((0, ncomp), dtype=np.float32)
                write_dataset(g, name, data)
        stacey = fresh_dir(os.path.join(fresh_dir(os.path.join(base, "Boundary")), "Stacey"))
        write_dataset(stacey, "IndexMapping", stacey_index_mapping(engine))
        write_dataset(stacey, "ElasticAcceleration", engine.stacey_record())
        write_dataset(stacey, "AcousticAcceleration", np.zeros((0, 5, 5, engine.nstep, 1), dtype=np.float32))
        write_dataset(stacey, "PoroelasticAcceleration", np.zeros((0, 5, 5, engine.nstep, 4), dtype=np.float32))

    def finalize(self, engine):
        pass


class WavefieldReader:
    """Periodic task of a combined run reading ``<folder>/ForwardWavefield``
    step ``istep`` into the buffer wavefield and the Stacey record
    (periodic_tasks/wavefield_reader.hpp, io/wavefield/reader.tpp)."""

    def __init__(self, folder: str, time_interval: int = 0, include_last_step: bool = True):
        self.root = os.path.join(folder, "ForwardWavefield")
        self.time_interval, self.include_last_step = int(time_interval), bool(include_last_step)

    should_run = WavefieldWriter.should_run

    def initialize(self, engine):
        pass

    def run(self, engine, istep: int):
        base = os.path.join(self.root, f"Step{istep:06d}")
        g = os.path.join(base, "elastic_psv")
        for name, which in zip(("Displacement", "Velocity", "Acceleration"),
                               ("displacement", "velocity", "acceleration")):
            data = read_dataset(g, name, np.float32, shape=(engine.nglob, engine.ncomp))
            engine.set_wavefield_field("buffer", which, data.T)
        for medium in STEP_MEDIA:
            if medium != "elastic_psv":
                for name in (("Potential",) if medium == "acoustic" else ("Displacement",)):
                    shape = read_meta(os.path.join(base, medium), name)
                    if shape[0] != 0:
                        raise NotImplementedError(f"{medium} wavefields: the engine's meshes are elastic P-SV")
        stacey = os.path.join(base, "Boundary", "Stacey")
        mapping = read_dataset(stacey, "IndexMapping", np.int32, shape=(engine.nspec,))
        if not np.array_equal(mapping, stacey_index_mapping(engine)):
            raise ValueError("the Stacey elements of the file differ from this mesh's")
        nstacey = engine.stacey_elements()
        rec = read_dataset(stacey, "ElasticAcceleration", np.float32, shape=(nstacey, 5, 5, engine.nstep, 2))
        engine.set_stacey_record(rec)

    def finalize(self, engine):
        pass


# ---- kernels ----------------------------------------------------------------
def write_kernels(folder: str, engine) -> str:
    """``<folder>/Kernels/elastic_psv_isotropic/{X, Z, rho, mu, kappa, rhop,
    alpha, beta}`` (kernel_writer::write)."""
    root = fresh_dir(os.path.join(folder, "Kernels"))
    g = fresh_dir(os.path.join(root, "elastic_psv_isotropic"))
    coords, _, _ = _compute_order_fields(engine)
    c2m = engine.compute_to_mesh()
    K = engine.kernels()  # mesh order
    # DomainView with one-element chunks: index = e * 25 + iz + 5 ix, i.e. the
    # (n, 5, 5) array transposed per element and written row-major
    domain = lambda a: np.ascontiguousarray(np.transpose(a, (0, 2, 1)))  # noqa: E731
    write_dataset(g, "X", domain(coords[..., 0]).reshape(-1, 5, 5), order="C")
    write_dataset(g, "Z", domain(coords[..., 1]).reshape(-1, 5, 5), order="C")
    for name in KERNEL_NAMES:
        write_dataset(g, name, domain(K[name][c2m]), order="C")
    return root


def read_kernels(folder: str, medium_property: str = "elastic_psv_isotropic"):
    """dict name -> (n, 5, 5) [e, iz, ix] of a Kernels directory."""
    g = os.path.join(folder, "Kernels", medium_property)
    out = {}
    for name in ("X", "Z") + KERNEL_NAMES:
        a = read_dataset(g, name, np.float64, order="C")
        out[name] = np.transpose(a, (0, 2, 1)) if a.ndim == 3 else a
    return out
