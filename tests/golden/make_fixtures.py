"""Write the reference's own test data for the 2-D elastic path as compact
fixtures under tests/golden/.

Run once in a container that has the reference checkout:

    python tests/golden/make_fixtures.py /root/reference

Only data files are taken (mesh databases, SPECFEM2D traces, index mapping and
mass-matrix arrays, station and source parameters, and the text run inputs:
specfem_config.yaml, sources.yaml, STATIONS, topography.dat and the source-reader test data).  They come from
PrincetonUniversity/SPECFEMPP ``tests/unit-tests`` (GPL-3.0) and are stored
re-encoded (gzip / npz) with their provenance recorded in ``MANIFEST.json``.
"""
from __future__ import annotations

import gzip
import json
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from specfempp_amd.io.fortran_db import read_fortran_array, read_stations, read_trace  # noqa: E402

UT = "tests/unit-tests"
NEWMARK = "displacement_tests/Newmark/serial"
_RICKER4 = [dict(x=x, z=z, angle=0.0, f0=10.0, factor=1e10, tshift=0.0)
            for (x, z) in [(500.0, 500.0), (3500.0, 500.0), (500.0, 2500.0), (3500.0, 2500.0)]]
CASES = {
    # case name: (fixture dir under tests/unit-tests, seismogram kinds (d, v, a,
    # p = pressure), sources, dt, nstep, source file)
    "elastic_free_80x60": (
        f"{NEWMARK}/HomogeneousElasticDomain", ["d", "v", "a"],
        [dict(x=2500.0, z=2500.0, angle=0.0, f0=10.0, factor=1e10, tshift=0.0)], 1.1e-3, 300, "sources.yaml"),
    "elastic_stacey_80x60": (
        f"{NEWMARK}/HomogeneousElasticDomainStaceyDirichletBC", ["d"], _RICKER4, 1.1e-3, 300, "sources.yaml"),
    # acoustic medium: free surface on all four sides, a Dirac force source
    "acoustic_free_80x60": (
        f"{NEWMARK}/HomogeneousAcousticDomain", ["d", "v", "a", "p"],
        [dict(x=2500.0, z=2500.0, angle=0.0, factor=1e10, tshift=0.0, stf="dirac")], 1.1e-3, 300,
        "sources.yaml"),
    "acoustic_dirichlet_80x60": (
        f"{NEWMARK}/HomogeneousAcousticDomainDirichletBC", ["d"], _RICKER4, 1.1e-3, 300, "sources.yaml"),
    # Stacey on bottom, left and right, free surface on top
    "acoustic_composite_80x60": (
        f"{NEWMARK}/HomogeneousAcousticDomainCompositeStaceyDirichletBC", ["d"], _RICKER4, 1.1e-3, 300,
        "source.yaml"),
    # acoustic medium with Stacey on all four sides (the reference's
    # HomogeneousElasticDomainStaceyBC case holds an acoustic material, vs = 0)
    "acoustic_stacey_80x60": (
        f"{NEWMARK}/HomogeneousElasticDomainStaceyBC", ["d"], _RICKER4, 1.1e-3, 300, "sources.yaml"),
    # elastic below, acoustic above one horizontal interface (9-node elements)
    "coupled_acoustic_top": (
        f"{NEWMARK}/AcousticElasticCoupledDomain", ["d"],
        [dict(x=1575.0, z=2900.0, angle=0.0, f0=10.0, factor=1e9, tshift=0.0)], 0.85e-3, 600, "sources.yaml"),
    # acoustic below, elastic above (4-node elements)
    "coupled_acoustic_bottom": (
        f"{NEWMARK}/AcousticElasticCoupledDomainVariant2", ["d"],
        [dict(x=1575.0, z=2900.0, angle=0.0, f0=10.0, factor=1e9, tshift=0.0)], 0.85e-3, 600, "sources.yaml"),
}


MESH_CASES = {
    "flat_topography_80x60": "data/mesh/simple_mesh_flat_topography",
    "curved_topography_80x60": "data/mesh/simple_mesh_curved_topography",
}


def gz_copy(src: str, dst: str) -> None:
    with open(src, "rb") as fi, open(dst, "wb") as raw:
        with gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as fo:
            fo.write(fi.read())


def main(ref: str) -> None:
    manifest = {}
    for name, (rel, kinds, sources, dt, nstep, src_file) in CASES.items():
        src_dir = os.path.join(ref, UT, rel)
        out = os.path.join(HERE, name)
        os.makedirs(out, exist_ok=True)
        gz_copy(os.path.join(src_dir, "database.bin"), os.path.join(out, "database.bin.gz"))
        stations = read_stations(os.path.join(src_dir, "STATIONS"))
        traces = {}
        time = None
        for k in kinds:
            rows = []
            for (sta, net, _x, _z) in stations:
                comps = []
                for comp in (("PRE",) if k == "p" else ("BXX", "BXZ")):
                    tr = read_trace(os.path.join(src_dir, "traces", f"{net}.{sta}.S2.{comp}.sem{k}"))
                    if time is None:
                        time = tr[:, 0]
                    comps.append(tr[:, 1])
                rows.append(comps)
            traces[k] = np.asarray(rows)  # (nrec, 2, nsamples); pressure (nrec, 1, nsamples)
        np.savez_compressed(os.path.join(out, "traces.npz"), time=time, **traces)
        # the run inputs as the reference's test holds them (text, unchanged)
        for fname in ("specfem_config.yaml", src_file, "STATIONS"):
            shutil.copyfile(os.path.join(src_dir, fname), os.path.join(out, fname))
        with open(os.path.join(out, "case.json"), "w") as fh:
            json.dump({
                "source_dir": f"{UT}/{rel}",
                "dt": dt, "nstep": nstep, "nstep_between_samples": 1,
                "stations": [dict(name=s, network=n, x=x, z=z) for (s, n, x, z) in stations],
                "sources": sources, "trace_kinds": kinds, "tolerance": 1e-3,
            }, fh, indent=1)
        manifest[name] = f"{UT}/{rel}"

    # Mass-matrix fixture (tests/unit-tests/domain/serial/test1).
    rel = "domain/serial/test1"
    src_dir = os.path.join(ref, UT, rel)
    out = os.path.join(HERE, "mass_80x60")
    os.makedirs(out, exist_ok=True)
    gz_copy(os.path.join(src_dir, "database.bin"), os.path.join(out, "database.bin.gz"))
    ibool = read_fortran_array(os.path.join(src_dir, "index_mapping.bin"), "<i4")
    rmass = read_fortran_array(os.path.join(src_dir, "rmass_inverse_elastic.bin"), "<f8")
    np.savez_compressed(os.path.join(out, "mass.npz"), index_mapping=ibool, rmass_inverse=rmass)
    with open(os.path.join(out, "case.json"), "w") as fh:
        json.dump({"source_dir": f"{UT}/{rel}", "dt": 1.1e-3, "tolerance": 1e-5,
                   "index_mapping_layout": "(nspec, ngllz, ngllx) row-major, 0-based, mesh order",
                   "rmass_layout": "(nglob, 2) row-major"}, fh, indent=1)
    manifest["mass_80x60"] = f"{UT}/{rel}"

    # Mesh-reader data (tests/unit-tests/data/mesh): two 80 x 60 meshes of
    # 9-node elements under flat and curved topography, with the material
    # properties the reader test expects (tests/unit-tests/mesh/materials/
    # properties.cpp:15-35: kappa, mu, rho).  The run parameters (dt, nstep,
    # the force source of sources.yaml) are for the engine-vs-oracle runs.
    for name, rel in MESH_CASES.items():
        src_dir = os.path.join(ref, UT, rel)
        out = os.path.join(HERE, name)
        os.makedirs(out, exist_ok=True)
        gz_copy(os.path.join(src_dir, "database.bin"), os.path.join(out, "database.bin.gz"))
        for fname in ("sources.yaml", "STATIONS", "topography.dat"):
            shutil.copyfile(os.path.join(src_dir, fname), os.path.join(out, fname))
        stations = read_stations(os.path.join(src_dir, "STATIONS"))
        with open(os.path.join(out, "case.json"), "w") as fh:
            json.dump({
                "source_dir": f"{UT}/{rel}",
                "properties_known_answer": {"kappa": 13499997600.24, "mu": 8100001799.82, "rho": 2700.0,
                                            "from": f"{UT}/mesh/materials/properties.cpp:15-35"},
                "nspec": 4800, "ngnod": 9,
                "dt": 8e-4, "nstep": 1500, "nstep_between_samples": 1,
                "stations": [dict(name=s, network=n, x=x, z=z) for (s, n, x, z) in stations],
                "sources": [dict(x=2500.0, z=2500.0, angle=0.0, f0=10.0, factor=1e10, tshift=0.0)],
            }, fh, indent=1)
        manifest[name] = f"{UT}/{rel}"

    # Geometry and point-location data: the database shared by the
    # partial-derivative test and the locate_point test (the two files are
    # byte-identical), SPECFEM2D's xix, xiz, gammax, gammaz and jacobian
    # ((nspec, ngllz, ngllx) row-major, mesh order; compute_partial_derivatives_tests.cpp)
    # and the locate_point known answers (algorithms/locate.cpp:37-58: mesh
    # element, xi, gamma of five points; ispec within 0, xi and gamma within
    # 1e-4, and the inverse map within 1e-2 m).
    rel = "compute/partial_derivatives/serial"
    src_dir = os.path.join(ref, UT, rel)
    out = os.path.join(HERE, "geometry_80x60")
    os.makedirs(out, exist_ok=True)
    db = os.path.join(src_dir, "Database00000.bin")
    with open(db, "rb") as f1, open(os.path.join(ref, UT, "algorithms/serial/database.bin"), "rb") as f2:
        assert f1.read() == f2.read(), "the two databases differ"
    gz_copy(db, os.path.join(out, "database.bin.gz"))
    arrays = {k: read_fortran_array(os.path.join(src_dir, f"{k}_00000.bin"), "<f8").reshape(-1, 5, 5)
              for k in ("xix", "xiz", "gammax", "gammaz", "jacobian")}
    np.savez_compressed(os.path.join(out, "partial_derivatives.npz"), **arrays)
    with open(os.path.join(out, "case.json"), "w") as fh:
        json.dump({"source_dir": f"{UT}/{rel}", "also": f"{UT}/algorithms/serial/database.bin",
                   "partial_derivatives_layout": "(nspec, ngllz, ngllx) row-major, mesh order",
                   "partial_derivatives_tolerance": 1e-2,
                   "locate_known_answers": {
                       "from": f"{UT}/algorithms/locate.cpp:37-58",
                       "points": [[606.313, 957.341], [1835.050, 146.444], [1539.578, 242.743],
                                  [3497.449, 2865.483], [3018.139, 530.283]],
                       "mesh_ispec": [1452, 196, 350, 4549, 700],
                       "xi": [-0.7475, 0.4020, 0.5831, 0.8980, -0.2744],
                       "gamma": [0.2905, 0.05832, -0.2247, 0.4248, 0.3945],
                       "tolerance_local": 1e-4, "tolerance_global": 1e-2}}, fh, indent=1)
    manifest["geometry_80x60"] = f"{UT}/{rel} + {UT}/algorithms/serial"

    # source-file reader data (tests/unit-tests/io/sources)
    rel = "io/sources"
    out = os.path.join(HERE, "io_sources")
    os.makedirs(out, exist_ok=True)
    for fname in ("single_force.yaml", "single_moment_tensor.yaml", "single_cosserat_force.yaml"):
        shutil.copyfile(os.path.join(ref, UT, rel, "data", fname), os.path.join(out, fname))
    manifest["io_sources"] = f"{UT}/{rel}/data"

    with open(os.path.join(HERE, "MANIFEST.json"), "w") as fh:
        json.dump({"origin": "PrincetonUniversity/SPECFEMPP tests/unit-tests (GPL-3.0)",
                   "cases": manifest}, fh, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
