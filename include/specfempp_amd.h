/*
 * specfempp_amd — MI355X-native spectral-element Newmark engine, C ABI.
 *
 * This header is the drop-in boundary for the SPECFEM++ Newmark time loop.
 * Each entry point names the reference interface it replaces.  All functions
 * return 0 on success and a negative status on error; sfx_last_error() gives
 * the message of the last failing call on the calling thread.  No C++
 * exception crosses the boundary.  Host arrays passed in are borrowed for the
 * duration of the call only (the engine copies what it keeps).
 *
 * Threading: an engine is single-threaded (one host thread drives it); two
 * engines are independent.  All device work of an engine runs in order on the
 * engine's own HIP stream.
 */
#ifndef SPECFEMPP_AMD_H
#define SPECFEMPP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFX_OK 0
#define SFX_ERR_ARG -1
#define SFX_ERR_HIP -2
#define SFX_ERR_STATE -3
#define SFX_ERR_MESH -4

typedef struct sfx_engine sfx_engine;

/* Edge codes of an absorbing edge (read_boundaries.cpp:166-179). */
#define SFX_EDGE_BOTTOM 0
#define SFX_EDGE_RIGHT 1
#define SFX_EDGE_TOP 2
#define SFX_EDGE_LEFT 3

/* Corner codes of an acoustic free-surface entry that names one control node
 * (read_boundaries.cpp:218-232). */
#define SFX_CORNER_BOTTOM_LEFT 4
#define SFX_CORNER_BOTTOM_RIGHT 5
#define SFX_CORNER_TOP_RIGHT 6
#define SFX_CORNER_TOP_LEFT 7

/* 2-D mesh: the content of specfem::mesh::mesh<dim2> for elastic P-SV and
 * acoustic isotropic domains (include/mesh/dim2/mesh.hpp; read by
 * src/io/mesh/impl/fortran/dim2/mesh.cpp:26-193).  A material with vs = 0 is
 * acoustic (read_material_properties.cpp:172-238). */
typedef struct {
  int nspec;            /* spectral elements */
  int ngnod;            /* control nodes per element: 4 or 9 */
  int npgeo;            /* control nodes */
  const double *coorg;  /* (npgeo, 2): x, z */
  const int *knods;     /* (nspec, ngnod): 0-based control node of each element */
  const int *material;  /* (nspec): 0-based material of each element */
  int nmaterials;
  const double *rho, *vp, *vs; /* per material (kg/m^3, m/s, m/s) */
  int nabs;             /* absorbing (Stacey) edges */
  const int *abs_ispec; /* (nabs): mesh element of each edge */
  const int *abs_edge;  /* (nabs): SFX_EDGE_* */
  /* acoustic free-surface (Dirichlet) entries, specfem::mesh::
   * acoustic_free_surface (read_boundaries.cpp:206-268): element and
   * SFX_EDGE_* or SFX_CORNER_* */
  int nafs;
  const int *afs_ispec;
  const int *afs_type;
  /* fluid-solid interface edges, specfem::mesh::coupled_interfaces
   * (read_interfaces.cpp:9-30): the acoustic and the elastic mesh element of
   * each shared edge */
  int nfs;
  const int *fs_acoustic;
  const int *fs_elastic;
} sfx_mesh2d;

/* 3-D structured hexahedral mesh (NGLL = 5).  Elements are numbered x
 * fastest, then y, then z; z points up.  The reference holds no 3-D solver:
 * this is the build's own mesh model, following the SPECFEM3D database
 * semantics read by src/io/mesh/impl/fortran/dim3/mesh.cpp:43-797 (ibool,
 * regular-element geometry, per-point kappa/mu/rho, Stacey faces). */
typedef struct {
  int nx, ny, nz;
  double x0, y0, z0;    /* origin (minimum corner) */
  double hx, hy, hz;    /* element size */
  const int *material;  /* (nz, ny, nx): 0-based material of each element */
  int nmaterials;
  const double *rho, *vp, *vs;
  int absorbing[6];     /* Stacey on faces xmin, xmax, ymin, ymax, zmin, zmax (1 = yes) */
} sfx_mesh3d;

/* 3-D mesh database: the elastic content of specfem::mesh::mesh<dim3> as
 * src/io/mesh/impl/fortran/dim3/mesh.cpp:43-797 reads it (generate_databases
 * output): numbering, GLL point coordinates, the regular/irregular geometry
 * split (mesh.cpp:150-151, 208-265), per-point kappa, mu, rho, rho vp, rho vs
 * and the absorbing faces.  Element arrays are (nspec, 5, 5, 5) with the x
 * index fastest (the record order of the reference's (nspec, ngllx, nglly,
 * ngllz) views); indices are 0-based.  Elements not listed as irregular use
 * xi_x = eta_y = gamma_z = xix_regular and jacobian_regular; the per-point
 * derivatives are read for the irregular ones. */
typedef struct {
  int nspec, nglob;
  const int *ibool;                     /* (nspec, 125) */
  const float *x, *y, *z;               /* (nglob) */
  int nspec_irregular;
  const int *irregular_element_number;  /* (nspec_irregular) */
  float xix_regular, jacobian_regular;
  const float *xix, *xiy, *xiz, *etax, *etay, *etaz, *gammax, *gammay, *gammaz, *jacobian; /* (nspec, 125) */
  const float *kappa, *mu, *rho;        /* (nspec, 125) */
  const float *rho_vp, *rho_vs;         /* (nspec, 125): the Stacey impedances */
  const int *ispec_type;                /* (nspec): 0 acoustic, 1 elastic, 2 poroelastic; elastic served */
  int nabs;                             /* absorbing faces */
  const int *abs_ispec;                 /* (nabs) */
  const int *abs_ijk;                   /* (nabs, 3, 25): i, j, k of every face point, point fastest */
  const float *abs_jacobian2Dw;         /* (nabs, 25): surface jacobian times the in-face GLL weights */
  const float *abs_normal;              /* (nabs, 3, 25): unit outward normal, point fastest */
} sfx_mesh3d_db;

/* Source kinds (src/source/force_source.cpp, moment_tensor_source.cpp). */
#define SFX_SOURCE_FORCE 0
#define SFX_SOURCE_MOMENT_TENSOR 1

/* Source time functions (src/source_time_function/ricker.cpp, dirac.cpp).
 * The Dirac is a narrow Gaussian of frequency 1 / (10 dt); f0 is unused. */
#define SFX_STF_RICKER 0
#define SFX_STF_DIRAC 1

/* One source.  2-D: position (x, z), force angle in degrees, moment tensor
 * (m[0]=Mxx, m[1]=Mzz, m[2]=Mxz); a force source in an acoustic element is a
 * pressure source (force_source.cpp:51-56).  3-D: position (x, y, z), force
 * vector (fx, fy, fz), moment tensor m = (Mxx, Myy, Mzz, Mxy, Mxz, Myz). */
typedef struct {
  int kind;
  double x, y, z;
  double angle;
  double fx, fy, fz;
  double m[6];
  double f0, tshift, factor;
  int stf; /* SFX_STF_* */
} sfx_source;

typedef struct {
  double x, y, z;
  double angle; /* 2-D station rotation in degrees (receivers.hpp:126-141) */
} sfx_receiver;

/* Seismogram kinds (specfem::wavefield::type). */
#define SFX_DISPLACEMENT 0
#define SFX_VELOCITY 1
#define SFX_ACCELERATION 2
/* -(lambda + 2/3 mu) div u at the receiver (wavefield.hpp:49-85), 2-D; its
 * second component is 0. */
#define SFX_PRESSURE 3

typedef struct {
  double dt;
  int nstep;
  int nstep_between_samples;
  int nseismo_types;
  int seismo_types[4]; /* up to the four kinds below (the reference records any subset) */
  int chunk_size;   /* storage chunk size of the numbering (chunk_config.hpp:12-28); 1 = CPU no-SIMD */
  int scatter;      /* weak-form scatter: SFX_SCATTER_COLORED, _ATOMIC, _GATHER or _BLOCK (3-D) */
  double start_time; /* 0: the earliest source onset (automatic); otherwise the user start time, the
                      * negated time-scheme "t0" of the configuration (time_scheme.cpp:46-50), which
                      * keeps every source's own tshift (src/io/sources.cpp:141-155) */
} sfx_run_desc;

/* Scatter of element forces into the global acceleration.  COLORED groups the
 * elements into colors whose members share no global point and adds with plain
 * read-modify-write, one launch per color; ATOMIC adds every element point with
 * a float atomic in one launch (the reference's atomic_add_on_device,
 * include/compute/fields/data_access.tpp:683-720). */
#define SFX_SCATTER_COLORED 0
#define SFX_SCATTER_ATOMIC 1
/* GATHER (3-D structured meshes): the stiffness kernel stores each element's
 * forces with coalesced writes; the fused update then sums, for every global
 * point, the contributions of its elements in a fixed order (deterministic,
 * no atomics, the acceleration is never stored between the two). */
#define SFX_SCATTER_GATHER 2
/* BLOCK (3-D structured meshes with even nx, ny, nz): one workgroup computes
 * 2 x 2 x 2 elements, sums their forces in on-chip memory in a fixed order and
 * updates the points inside the block directly; only the forces on block
 * faces pass through device memory to a second, small update kernel
 * (deterministic, no atomics). */
#define SFX_SCATTER_BLOCK 3

/* ---- library ---------------------------------------------------------- */
int sfx_version(void);
const char *sfx_last_error(void);

/* ---- engine lifetime ------------------------------------------------------
 * Replaces the construction of specfem::compute::assembly
 * (include/compute/assembly/assembly.hpp:34-119; src/compute/assembly/assembly.cpp:6-101)
 * followed by time_scheme::link_assembly (include/timescheme/newmark.hpp:103-105).
 * Builds numbering, geometry, properties, boundaries, sources and receivers on
 * the host and copies them to the current HIP device. */
int sfx_engine_create_2d(const sfx_mesh2d *mesh, int nsources, const sfx_source *sources, int nreceivers,
                         const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out);
int sfx_engine_create_3d(const sfx_mesh3d *mesh, int nsources, const sfx_source *sources, int nreceivers,
                         const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out);
/* An engine on a 3-D mesh database (general hexahedra, per-point geometry on
 * irregular elements): SFX_SCATTER_COLORED (greedy element coloring) or
 * _ATOMIC.  Sources and receivers are located on the GLL points' isoparametric
 * map (nearest point, then Newton iterations in the elements that hold it). */
int sfx_engine_create_3d_db(const sfx_mesh3d_db *mesh, int nsources, const sfx_source *sources, int nreceivers,
                            const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out);
int sfx_engine_create_host_3d_db(const sfx_mesh3d_db *mesh, int nsources, const sfx_source *sources,
                                 int nreceivers, const sfx_receiver *receivers, const sfx_run_desc *run,
                                 sfx_engine **out);
/* Host-only engines: the host half of the construction above (the
 * specfem::compute::assembly host views: numbering compute_mesh.cpp:45-170,
 * ordering :240-279, locate_point.cpp, sources and receivers) and no device
 * work at all.  The data-contract getters below (sizes, t0, index mapping,
 * field index, compute order, locations) and sfx_batch_create accept them;
 * every other call returns an error. */
int sfx_engine_create_host_2d(const sfx_mesh2d *mesh, int nsources, const sfx_source *sources, int nreceivers,
                              const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out);
int sfx_engine_create_host_3d(const sfx_mesh3d *mesh, int nsources, const sfx_source *sources, int nreceivers,
                              const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out);
void sfx_engine_destroy(sfx_engine *engine);

/* Sizes: dim, nspec, nglob, ngll, nsig (seismogram samples), nrec, nsrc. */
int sfx_get_sizes(const sfx_engine *engine, int64_t *sizes7);
/* Storage layout the engine chose (no counterpart in the reference; for
 * byte models and tests): [0] geometry of the element forces: 0 per-point
 * arrays, 1 one record per element (every element's geometry is the same at
 * all of its points), 4 or 9 recomputed at every point from the element's 4
 * or 9 control nodes (2-D gather path; the same bits as the per-point
 * arrays); [1] 1 if the 2-D gather path reads kappa and mu once per element
 * (both the same at all points of every element), else 0; [2] the scatter
 * mode; [3] 1 if the device numbers the points in coordinate row order (2-D
 * gather path), else 0. */
int sfx_get_layout(const sfx_engine *engine, int64_t *layout4);
/* Per-medium sizes of a 2-D mesh (specfem::compute::simulation_field's
 * per-medium fields, field_impl.tpp:23-103): elastic elements, acoustic
 * elements, points of the elastic field, points of the acoustic field.  A mesh
 * without acoustic elements gives (nspec, 0, nglob, 0). */
int sfx_get_media_sizes(const sfx_engine *engine, int64_t *sizes4);
/* Acoustic field index of every global point, -1 where no acoustic element
 * touches it (field_impl.tpp:39-62 for the acoustic medium).  2-D. */
int sfx_get_field_index_acoustic(const sfx_engine *engine, int32_t *out);
/* Simulation start time t0 (src/io/sources.cpp:128-158). */
int sfx_get_t0(const sfx_engine *engine, double *t0);

/* ---- data contract (bit-exact comparison against the reference) ------- */
/* index_mapping (nspec, ngll, ngll[, ngll]) in compute order
 * (specfem::compute::points::h_index_mapping, compute_mesh.cpp:45-170). */
int sfx_get_index_mapping(const sfx_engine *engine, int32_t *out);
/* per-medium numbering iglob -> field index (field_impl.tpp:39-62). */
int sfx_get_field_index(const sfx_engine *engine, int32_t *out);
/* compute order -> mesh order (mesh_to_compute_mapping, compute_mesh.cpp:240-279). */
int sfx_get_compute_to_mesh(const sfx_engine *engine, int32_t *out);
/* element and reference coordinates of every source / receiver
 * (locate_point, src/algorithms/locate_point.cpp:120-176). */
int sfx_get_locations(const sfx_engine *engine, int which /*0 sources, 1 receivers*/, int32_t *ispec,
                      double *xi_eta_gamma /* (n, dim) */);

/* ---- time-scheme seam (drop-in for time_scheme / domain_kernels) ---------
 * sfx_initialize: kernels.initialize(dt) — mass matrix and its inverse
 *   (domain_kernels.hpp:136-163); also zeroes the wavefield.
 * sfx_predictor: apply_predictor_phase_forward (newmark.tpp:140-159); returns
 *   the number of degrees of freedom updated (nglob * ncomp).
 * sfx_update_wavefields: update_wavefields<elastic> (domain_kernels.hpp:69-126):
 *   sources, stiffness (+ Stacey), mass division; returns elements updated.
 * sfx_corrector: apply_corrector_phase_forward (newmark.tpp:119-137).
 * sfx_compute_seismograms: compute_seismograms(isig) (domain_kernels.hpp:172-186). */
int sfx_initialize(sfx_engine *engine);
int sfx_predictor(sfx_engine *engine);
int sfx_update_wavefields(sfx_engine *engine, int istep);
int sfx_corrector(sfx_engine *engine);
int sfx_compute_seismograms(sfx_engine *engine, int isig);

/* ---- wavefield, medium and direction seams -------------------------------------
 * The calls above act on the forward wavefield of the elastic medium.  These
 * take the three selectors the reference's interfaces carry, so that an adapter
 * can implement every pure virtual of specfem::time_scheme::time_scheme
 * (include/timescheme/timescheme.hpp:173-207) and drive
 * time_marching<combined>::run (include/solver/time_marching.tpp:126-247)
 * seam by seam:
 *   medium     specfem::element::medium_tag.  Phases of a medium with no
 *              element in the mesh update nothing and return 0, as the
 *              reference's do (newmark.tpp:11-116 loops over an empty field).
 *   wavefield  specfem::wavefield::simulation_field: forward, adjoint,
 *              backward, buffer (the forward state read back,
 *              time_marching.tpp:201-209).  2-D engines hold all four (the
 *              last three are allocated at first use and share the forward
 *              field's mass matrix); 3-D engines hold the forward field.
 *   direction  newmark<forward>/<combined> phases: forward steps with +dt,
 *              backward with -dt (newmark.tpp:161-242).
 * sfx_predictor_phase / sfx_corrector_phase: apply_{predictor,corrector}_phase_
 *   {forward,backward}(medium) (newmark.tpp:119-242); returns the medium's
 *   degrees of freedom (nglob * ncomp, or 0).
 * sfx_update_wavefields_medium: domain_kernels<wavefield>::update_wavefields
 *   <medium>(istep) (domain_kernels.hpp:69-126); returns the medium's elements.
 *   The backward field replays the Stacey record of a completed
 *   sfx_forward_record (compute_stiffness_interaction.tpp:119-136).
 * sfx_compute_seismograms_wavefield: domain_kernels<wavefield>::compute_seismograms.
 * sfx_initialize_wavefield: domain_kernels<wavefield>::initialize(dt): zeroes
 *   the field (the forward one as sfx_initialize); initializing the adjoint
 *   field also zeroes the Fréchet kernels, which the reference's assembly
 *   creates at zero.
 * sfx_copy_wavefield: specfem::compute::deep_copy(dst, src) (u, v, a).
 * sfx_compute_derivatives: frechet_kernels::compute_derivatives(dt)
 *   (include/kokkos_kernels/frechet_kernels.hpp:19-53,
 *   compute_material_derivatives.tpp:17-154) from the adjoint and backward
 *   fields; sfx_get_kernels returns the accumulated kernels.  2-D.
 * sfx_get_wavefield / sfx_set_wavefield: sfx_get_field / sfx_set_field of any
 *   wavefield. */
#define SFX_MEDIUM_ELASTIC_PSV 0
#define SFX_MEDIUM_ACOUSTIC 1
#define SFX_MEDIUM_ELASTIC_SH 2
#define SFX_MEDIUM_POROELASTIC 3
#define SFX_MEDIUM_ELASTIC_PSV_T 4
#define SFX_WAVEFIELD_FORWARD 0
#define SFX_WAVEFIELD_ADJOINT 1
#define SFX_WAVEFIELD_BACKWARD 2
#define SFX_WAVEFIELD_BUFFER 3
#define SFX_DIRECTION_FORWARD 0
#define SFX_DIRECTION_BACKWARD 1
int sfx_initialize_wavefield(sfx_engine *engine, int wavefield);
int sfx_predictor_phase(sfx_engine *engine, int medium, int wavefield, int direction);
int sfx_update_wavefields_medium(sfx_engine *engine, int medium, int wavefield, int istep);
int sfx_corrector_phase(sfx_engine *engine, int medium, int wavefield, int direction);
int sfx_compute_seismograms_wavefield(sfx_engine *engine, int wavefield, int isig);
int sfx_copy_wavefield(sfx_engine *engine, int dst_wavefield, int src_wavefield);
int sfx_compute_derivatives(sfx_engine *engine, double dt);
int sfx_get_wavefield(const sfx_engine *engine, int wavefield, int which, float *out);
int sfx_set_wavefield(sfx_engine *engine, int wavefield, int which, const float *in);
/* The forward field of one medium: SFX_MEDIUM_ELASTIC_PSV gives (2, nglob_el)
 * as sfx_get_field; SFX_MEDIUM_ACOUSTIC gives the potential chi, chi_dot,
 * chi_dot_dot or the mass inverse, (nglob_ac), on the acoustic field index. */
int sfx_get_medium_field(const sfx_engine *engine, int medium, int which, float *out);
int sfx_set_medium_field(sfx_engine *engine, int medium, int which, const float *in);

/* ---- fused fast path --------------------------------------------------------
 * One Newmark step of time_marching<forward>::run (time_marching.tpp:33-109)
 * with the corrector of step istep, the seismogram sample and the predictor of
 * step istep+1 fused into one pass over the global points.
 * sfx_run: steps [first, first + count) — the whole loop when (0, nstep). */
int sfx_step(sfx_engine *engine, int istep);
int sfx_run(sfx_engine *engine, int first, int count);
int sfx_synchronize(sfx_engine *engine);

/* ---- run without allocation ---------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory or create HIP objects while stepping (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): sfx_run_noalloc runs
 * steps [first, first + count) of a 3-D SFX_SCATTER_GATHER engine exactly as
 * sfx_run does (same kernels, same bits), but the per-element forces live in a
 * static device array of SFX_NOALLOC_MAX_ELEMENTS x 384 floats (64 MiB per GPU,
 * held until the process ends) instead of an engine buffer, and no timing
 * event, graph or buffer is created.  Engines larger than the limit are
 * refused.  All engines on one GPU share that array: calls that overlap on one
 * GPU (two host threads) race.  The call returns with its work complete and
 * the stored acceleration assembled; sfx_get_timings is not updated. */
#define SFX_NOALLOC_MAX_ELEMENTS 43690
int sfx_run_noalloc(sfx_engine *engine, int first, int count);

/* ---- outputs ------------------------------------------------------------- */
/* seismograms (nsig, ntypes, nrec, ncomp) before station rotation
 * (receivers seismogram_components, compute_seismogram.tpp:103-135). */
int sfx_get_seismograms(const sfx_engine *engine, float *out);
/* which: 0 displacement, 1 velocity, 2 acceleration, 3 mass inverse;
 * out (ncomp, nglob) on the per-medium field index. */
int sfx_get_field(const sfx_engine *engine, int which, float *out);
int sfx_set_field(sfx_engine *engine, int which, const float *in);

/* ---- adjoint (2-D) ----------------------------------------------------------
 * sfx_forward_record: the whole forward loop (as sfx_run(0, nstep)) that also
 *   keeps what the combined run reads back: the local force of every point of
 *   every Stacey element at every step (the forward branch that stores
 *   boundary_values, compute_stiffness_interaction.tpp:251-255;
 *   boundary_medium_container.hpp:26-82) and the final corrected wavefield
 *   (the ForwardWavefield written by io/wavefield/writer.tpp).  Replaces a
 *   forward run with the wavefield writer task.  snapshot_every > 0 also keeps
 *   the corrected wavefield after steps 0, k, 2k, ... (< nstep - 1) in device
 *   memory; the combined run resets its backward field to them, bounding the
 *   drift of the time-reversed reconstruction (0: pure time reversal, as the
 *   reference).
 * sfx_set_adjoint_sources: adjoint sources (src/source/adjoint_source.cpp:9-90)
 *   at `positions` (x, z used) with external time functions
 *   (src/source_time_function/external.cpp) stf (nadj, nstep, 2), in forward
 *   time order.
 * sfx_adjoint_run: time_marching<combined>::run (time_marching.tpp:126-247):
 *   istep from nstep - 1 down to 0, the adjoint field forward in time, the
 *   backward field with -dt (newmark.tpp:161-242) replaying the recorded
 *   Stacey elements (compute_stiffness_interaction.tpp:119-136), and
 *   frechet_kernels::compute_derivatives (compute_material_derivatives.tpp)
 *   every step.  The seismogram buffer then holds the backward field's
 *   seismograms, latest sample first.  Consumes the recorded forward state.
 * sfx_get_kernels: kernels (6, nspec, ngll, ngll) in mesh element order, in the
 *   order rho, mu, kappa, rhop, alpha, beta (point::kernels,
 *   core/specfem/point/kernels.hpp:55-332; frechet_derivative.hpp:14-138). */
int sfx_forward_record(sfx_engine *engine, int snapshot_every);
/* Replace the per-point properties rho, kappa (bulk modulus), mu, each
 * (nspec, ngll, ngll) in mesh element order, as a model update of an inversion
 * does (the reference rebuilds compute::properties from a new model,
 * src/compute/compute_properties.cpp, with the mass matrix and the Stacey
 * impedances that depend on them).  2-D. */
int sfx_set_properties(sfx_engine *engine, const float *rho, const float *kappa, const float *mu);
int sfx_set_adjoint_sources(sfx_engine *engine, int nadj, const sfx_receiver *positions, const float *stf);
/* The Stacey record of the last sfx_forward_record, in the layout of the
 * reference's boundary_values.stacey.elastic values (nstacey, ngllz, ngllx,
 * nstep, 2) LayoutLeft, element fastest
 * (include/compute/boundary_values/impl/boundary_medium_container.hpp:26-40):
 * the local acceleration (stiffness + Stacey traction) of every point of every
 * Stacey element at every step, elements in compute order
 * (boundary_medium_container.tpp:20-31).  The wavefield writer saves it as
 * Boundary/Stacey/ElasticAcceleration (io/wavefield/writer.tpp:127-148).
 * sfx_set_stacey_record loads such a record (the wavefield reader,
 * io/wavefield/reader.tpp:44-62) and declares the forward wavefield to hold
 * the recorded run's final state, which the caller sets (sfx_set_wavefield or
 * sfx_copy_wavefield from the buffer) before sfx_adjoint_run; the backward
 * seam (sfx_update_wavefields_medium) replays it too.  2-D. */
int sfx_get_stacey_record(const sfx_engine *engine, float *out);
int sfx_set_stacey_record(sfx_engine *engine, const float *in);
/* GLL point coordinates (nspec, ngll, ngll, 2) in compute order, as the
 * reference stores them in type_real (mesh.points.h_coord,
 * compute_mesh.cpp:299-333); the wavefield and kernel writers save them.
 * 2-D. */
int sfx_get_point_coordinates(const sfx_engine *engine, float *out);
int sfx_adjoint_run(sfx_engine *engine);
int sfx_get_kernels(const sfx_engine *engine, float *out);
/* Integration weight w_i w_j J of every element point, (nspec, ngll, ngll) in
 * mesh element order (the factor of compute_mass_matrix.tpp:99-101): a kernel
 * density K at a point times this weight is the derivative of the misfit with
 * respect to the logarithm of that point's parameter.  2-D. */
int sfx_get_point_weights(const sfx_engine *engine, float *out);

/* ---- batched runs (2-D) -------------------------------------------------------
 * The reference runs one simulation per process (time_marching<forward>::run,
 * time_marching.tpp:9-124).  A batch advances many independent 2-D
 * simulations (meshes of any size, own sources, receivers, dt and step count)
 * in the same launches.  Engines are only read at creation (the batch copies
 * their data); they must record the same seismogram types.  Scatters:
 *   SFX_SCATTER_GATHER: one launch per time step covers every simulation still
 *     running (double-buffered state, no grid-wide synchronisation); each
 *     simulation gives its engine's SFX_SCATTER_GATHER results bit for bit;
 *   SFX_SCATTER_COLORED / _ATOMIC: one launch per kernel family per step
 *     (sources, stiffness per color, Stacey, seismograms, update).
 * Runs of steps between two simulation ends replay one captured HIP graph
 * when graphs are on.
 * sfx_batch_tune: times every scatter with 2, 5 and 10 elements per workgroup
 *   (50, 125 or 250 threads, one per element point), with and without graphs,
 *   on the batch itself and keeps the fastest; choice4 = (elements per
 *   workgroup, graphs on, colors, scatter).  sfx_batch_run starts from a zero
 *   state.
 * sfx_batch_set_config: the same choice made by the caller.
 * sfx_batch_get_seismograms: simulation i (creation order), laid out as
 *   sfx_get_seismograms of its engine.
 * sfx_batch_get_info: last run's device time (ms), elements per workgroup,
 *   graphs on, colors, padded points, total point-steps, scatter, launches
 *   per step. */
typedef struct sfx_batch sfx_batch;
int sfx_batch_create(int n, sfx_engine *const *engines, sfx_batch **out);
void sfx_batch_destroy(sfx_batch *batch);
int sfx_batch_tune(sfx_batch *batch, int steps_per_candidate, int *choice4);
int sfx_batch_set_config(sfx_batch *batch, int scatter, int elems_per_block, int use_graph);
int sfx_batch_run(sfx_batch *batch);
int sfx_batch_get_seismograms(const sfx_batch *batch, int i, float *out);
int sfx_batch_get_info(const sfx_batch *batch, double *out8);

/* ---- measurement -------------------------------------------------------------
 * Average device time (ms) of each kernel family over the last sfx_run, taken
 * with HIP events on the engine stream: out[0] stiffness, out[1] fused update,
 * out[2] boundary+source+seismogram, out[3] whole step. */
int sfx_get_timings(const sfx_engine *engine, double *out4);

#ifdef __cplusplus
}
#endif

#endif /* SPECFEMPP_AMD_H */
