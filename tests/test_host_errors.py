"""Error behaviour at the boundary (CPU, host-only engines): invalid inputs
are refused at creation with the library's message, as the reference throws
std::runtime_error from its setup (time_marching.tpp:90-108,
compute_stiffness_interaction.tpp:60-63).  Each case names the check it hits."""
import math

import numpy as np
import pytest

from specfempp_amd import Engine2D, Engine3D, SfxError, meshes


def _mesh2d(vs=1443.0):
    return meshes.quad_mesh(4, 3, 0.0, 200.0, 0.0, 150.0, materials=[meshes.elastic(2200.0, 2500.0, vs)])


def _mesh3d(n=2, material=None):
    mat = np.zeros((n, n, n), dtype=np.int32) if material is None else material
    return meshes.hex_mesh(n, n, n, 100.0, mat, [2500.0], [3000.0], [1730.0])


SRC2 = [dict(x=100.0, z=75.0, angle=0.0, f0=10.0, factor=1.0, tshift=0.0)]
REC2 = [dict(x=50.0, z=100.0)]
SRC3 = [dict(type="force", x=100.0, y=100.0, z=100.0, fx=0.0, fy=0.0, fz=1.0, f0=5.0, factor=1.0)]
REC3 = [dict(x=50.0, y=50.0, z=150.0)]


@pytest.mark.parametrize("kwargs,match", [
    (dict(dt=0.0), "dt must be positive"),
    (dict(nstep=-1), "nstep must be >= 0"),
    (dict(nsb=0), "nstep_between_samples must be >= 1"),
    (dict(start_time=math.inf), "start_time must be finite"),
    (dict(scatter="block"), "SFX_SCATTER_BLOCK is for 3-D meshes"),
])
def test_2d_run_description_is_checked(kwargs, match):
    args = dict(dt=1e-3, nstep=10, nsb=1, start_time=0.0, scatter="colored")
    args.update(kwargs)
    with pytest.raises(SfxError, match=match):
        Engine2D(_mesh2d(), SRC2, REC2, args["dt"], args["nstep"], args["nsb"], ("displacement",),
                 scatter=args["scatter"], start_time=args["start_time"], host_only=True)


def test_2d_inputs_are_checked():
    with pytest.raises(SfxError, match="vs >= 0"):
        Engine2D(_mesh2d(vs=-1.0), SRC2, REC2, 1e-3, 10, host_only=True)
    # vs = 0 is an acoustic medium: a moment tensor cannot act there
    with pytest.raises(SfxError, match="moment-tensor source in an acoustic element"):
        Engine2D(_mesh2d(vs=0.0), [dict(SRC2[0], type="moment_tensor", mxx=1.0)], REC2, 1e-3, 10, host_only=True)
    with pytest.raises(SfxError, match="SFX_SCATTER_COLORED or _ATOMIC"):
        Engine2D(_mesh2d(vs=0.0), SRC2, REC2, 1e-3, 10, scatter="gather", host_only=True)
    with pytest.raises(SfxError, match="source time function"):
        Engine2D(_mesh2d(), [dict(SRC2[0], stf="gaussian")], REC2, 1e-3, 10, host_only=True)
    with pytest.raises(SfxError, match="source f0 must be positive"):
        Engine2D(_mesh2d(), [dict(SRC2[0], f0=0.0)], REC2, 1e-3, 10, host_only=True)
    with pytest.raises(SfxError, match="non-finite source or receiver position"):
        Engine2D(_mesh2d(), SRC2, [dict(x=math.nan, z=10.0)], 1e-3, 10, host_only=True)
    with pytest.raises(ValueError, match="at most 4 seismogram types"):
        Engine2D(_mesh2d(), SRC2, REC2, 1e-3, 10, 1, ("displacement",) * 5, host_only=True)


def test_3d_inputs_are_checked():
    with pytest.raises(SfxError, match="pressure seismograms are 2-D only"):
        Engine3D(_mesh3d(), SRC3, REC3, 1e-3, 10, 1, ("pressure",), host_only=True)
    with pytest.raises(SfxError, match="material index out of range"):
        Engine3D(_mesh3d(material=np.full((2, 2, 2), 3, dtype=np.int32)), SRC3, REC3, 1e-3, 10, host_only=True)
    with pytest.raises(SfxError, match="empty 3-D mesh"):
        Engine3D(meshes.hex_mesh(0, 2, 2, 100.0, np.zeros((2, 2, 0), dtype=np.int32), [2500.0], [3000.0],
                                 [1730.0]), SRC3, REC3, 1e-3, 10, host_only=True)
    with pytest.raises(SfxError, match="needs even element counts"):
        Engine3D(meshes.hex_mesh(3, 2, 2, 100.0, np.zeros((2, 2, 3), dtype=np.int32), [2500.0], [3000.0],
                                 [1730.0]), SRC3, REC3, 1e-3, 10, scatter="block", host_only=True)
    # a valid host-only engine after the failures: the error state does not stick
    eng = Engine3D(_mesh3d(), SRC3, REC3, 1e-3, 10, host_only=True)
    assert eng.nglob == 9 ** 3
    eng.close()


def test_3d_mesh_beyond_32bit_device_indexing_is_refused():
    """224^3 elements: 3 x nglob_pad = 3 x 225^3 x 64 > 2^31 - 1.  The guard
    runs before any array of the mesh's size is allocated (assembly.cpp
    check_index_range_3d), so the refusal is immediate."""
    n = 224
    mesh = meshes.hex_mesh(n, n, n, 100.0, np.zeros((n, n, n), dtype=np.int32), [2500.0], [3000.0], [1730.0])
    with pytest.raises(SfxError, match="3-D mesh too large"):
        Engine3D(mesh, SRC3, REC3, 1e-3, 10, 1, ("displacement",), host_only=True)


def test_2d_mesh_beyond_32bit_device_indexing_is_refused():
    """nspec x 25 element points above 2^31 - 1 is refused before the element
    arrays are read."""
    mesh = _mesh2d()
    mesh.nspec = 90_000_000
    with pytest.raises(SfxError, match="2-D mesh too large"):
        Engine2D(mesh, SRC2, REC2, 1e-3, 10, 1, ("displacement",), host_only=True)
