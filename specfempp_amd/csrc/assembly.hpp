// Host-side assembly: everything the device kernels read, laid out for the
// GPU.  Built once per engine from the mesh description; the engine copies it
// to HBM and keeps only the small pieces it returns through the C ABI.
//
// This is the build's counterpart of specfem::compute::assembly
// (include/compute/assembly/assembly.hpp:34-119).  Layout differences from
// the reference (all internal):
//   - element arrays are element-major, point-minor ([e][p], p = (k,)j,i with i
//     fastest) so that a wavefront reads them with unit stride;
//   - the two-level index (index_mapping -> assembly_index_mapping,
//     include/compute/fields/simulation_field.hpp:134-145) is composed on the
//     host into one `gather` index per local point;
//   - Stacey boundaries are a flat list of (element point, normal, weight,
//     impedances) entries instead of per-element tags;
//   - fields are component-major (ncomp, nglob_pad) like the reference's
//     LayoutLeft (nglob, ncomp) views, padded to a multiple of 64 points;
//     3-D fields are stored in 4x4x4 tiles of lattice points.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/specfempp_amd.h"

namespace sfx {

struct HostAssembly {
  int dim = 2, ncomp = 2, npe = 25;      // npe: GLL points per element
  int nspec = 0, nglob = 0, nglob_pad = 0;
  int nspec_none = 0;                    // 2-D: elements [0, nspec_none) carry no Stacey edge
  // numbering
  std::vector<int32_t> index_mapping;    // (nspec, npe) reference iglob, compute order
  std::vector<int32_t> field_index;      // iglob -> per-medium field index
  std::vector<int32_t> compute_to_mesh;  // (nspec)
  std::vector<int32_t> gather;           // (nspec, npe) field index of each local point
  // quadrature (float, the reference's type_real)
  float xi[5], w[5], hprime[5][5];
  // 2-D geometry per local point (compute_partial_derivatives.cpp:46-91)
  std::vector<float> xix, xiz, gammax, gammaz, jac;
  // 2-D: when every element's geometry is the same at all of its points
  // (affine elements), one record per element (xix, xiz, gammax, gammaz, J,
  // padding to 8 floats) in compute order; empty otherwise
  std::vector<float> geo2d;
  // 3-D regular geometry (all elements are boxes of the same size)
  float xix3 = 0, ety3 = 0, gaz3 = 0, jac3 = 0;
  // 3-D database meshes (general hexahedra): per point the derivatives
  // xix, xiy, xiz, etax, etay, etaz, gammax, gammay, gammaz and the jacobian,
  // (nspec, 10, 125)
  bool general3d = false;
  std::vector<float> geo3;
  // properties per local point (kappa, mu; rho for mass and Stacey)
  std::vector<float> kappa, mu, rho;
  // Stacey entries: one per (element, local point) on an absorbing edge/face
  std::vector<int32_t> bnd_gather;
  std::vector<int32_t> bnd_slot;          // compute element * npe + local point of the entry
  std::vector<int32_t> bnd_seg;           // 3-D: entries [bnd_seg[s], bnd_seg[s+1]) share one slot
  std::vector<float> bnd_normal;         // (nbnd, dim) outward normal scaled as the reference stores it
  std::vector<float> bnd_weight;         // (nbnd) edge/face quadrature weight
  std::vector<float> bnd_rho_vp, bnd_rho_vs;
  // mass inverse (ncomp, nglob_pad) in device layout (2-D: field index; 3-D:
  // tiled lattice index); padding slots hold 0
  std::vector<float> mass_inverse;
  // sources
  int nsrc = 0;
  std::vector<int32_t> src_ispec;
  std::vector<float> src_array;          // (nsrc, ncomp, npe)
  std::vector<float> stf;                // (nstep, nsrc)
  std::vector<double> src_local;         // (nsrc, dim) reference coordinates
  std::vector<int32_t> src_order;        // application order (element type, then source)
  double t0 = 0;
  // receivers
  int nrec = 0;
  std::vector<int32_t> rec_ispec;
  std::vector<float> rec_h;              // (nrec, npe) interpolation weights
  std::vector<double> rec_local;         // (nrec, dim)
  // 3-D structured sizes; device fields in lattice order, lattice_of maps the
  // reference numbering (iglob) to the lattice index
  int nx = 0, ny = 0, nz = 0;
  int absorb_mask = 0;                   // bit f: face f (x-, x+, y-, y+, z-, z+) is absorbing
  int nint_pad = 0;                      // block layout: slots of the block-interior points
  std::vector<int32_t> lattice_of;
  // 2-D locator data (control nodes (nspec, ngnod, 2), GLL point coordinates)
  int ngnod2d = 0;
  std::vector<float> ctrl2d, px2d, pz2d;
  // 2-D: the shape-function derivatives the geometry above was computed from,
  // per GLL point (iz * 5 + ix) ngnod values in xi, then ngnod in gamma
  std::vector<float> dsh2d;
  // adjoint sources as single-component sources (see place_adjoint_sources_2d)
  int nadj = 0;
  std::vector<int32_t> adj_ispec, adj_order;
  std::vector<float> adj_array;          // (2 nadj, ncomp, npe)
  std::vector<float> adj_stf;            // (nstep, 2 nadj)
  // 2-D acoustic medium and elastic-acoustic coupling (meshes with vs = 0
  // materials).  Compute order puts the elastic elements first: [0, nspec_el)
  // elastic, [nspec_el, nspec) acoustic.  The elastic field has nglob_el
  // points (nglob_pad is its padded size), the acoustic field nglob_ac.
  int nspec_el = 0, nglob_el = 0;
  int nglob_ac = 0, nglob_ac_pad = 0;
  std::vector<int32_t> field_index_ac;   // iglob -> acoustic field index, or -1
  // per local point of every compute element (0 at elastic elements):
  std::vector<float> rho_inv;            // 1 / rho (acoustic properties, material.hpp:44-101)
  std::vector<float> ac_stacey;          // w |n| rho_vpinverse on acoustic Stacey edges, else 0
  std::vector<int32_t> ac_dirichlet;     // 1 on an acoustic free surface
  std::vector<float> mass_inverse_ac;    // (nglob_ac_pad), padding 0
  // Coupling, one entry per interface edge point.  cpl_ac: the acoustic
  // side's term factor (nx u_x + nz u_z) with n the acoustic normal, into the
  // acoustic field point `self` from the elastic field point `other`; cpl_el:
  // the elastic side's factor n chi_dot_dot into elastic point `self`.
  std::vector<int32_t> cpl_ac_self, cpl_ac_other, cpl_el_self, cpl_el_other;
  std::vector<float> cpl_ac_coef, cpl_el_coef;  // (n, 3): factor, nx, nz
  // sources and receivers by medium, in application order
  std::vector<int32_t> src_order_el, src_order_ac, rec_el, rec_ac;
  std::vector<int32_t> src_stf_kind;     // per source: SFX_STF_*
  int ncolors_el = 0;                    // color_offsets[0..ncolors_el] are the elastic colors
  bool coupled() const { return dim == 2 && nspec_el < nspec; }
  // Element colors for the scatter: device element slot s holds compute
  // element perm[s]; slots [color_offsets[c], color_offsets[c+1]) form color c,
  // whose elements share no global point.  One color when scatter is atomic.
  std::vector<int32_t> perm, inv_perm, color_offsets;
};

// Group elements into colors (no two elements of a color share a global point)
// and fill perm / inv_perm / color_offsets.  scatter = SFX_SCATTER_ATOMIC keeps
// compute order in a single color.
void color_elements(HostAssembly &a, int scatter);
// The same coloring, returned instead of stored.
struct ElementColoring {
  std::vector<int32_t> perm, inv_perm, color_offsets;
  int ncolors_el = 0;
};
ElementColoring element_coloring(const HostAssembly &a, int scatter);

// Replace the per-point properties ((nspec, 25), mesh element order) and
// rebuild the Stacey impedances and the mass matrix (2-D).
void set_properties_2d(HostAssembly &a, const float *rho, const float *kappa, const float *mu, double dt);
void mass_2d(HostAssembly &a, double dt);

// Locate adjoint sources (pos, nadj) with time functions stf (nadj, nstep, 2).
void place_adjoint_sources_2d(HostAssembly &a, int nadj, const sfx_receiver *pos, const float *stf, int nstep);

// Throw std::runtime_error / std::invalid_argument on bad input.
HostAssembly build_assembly_2d(const sfx_mesh2d &mesh, int nsrc, const sfx_source *src, int nrec,
                               const sfx_receiver *rec, const sfx_run_desc &run);
// Refuses 3-D meshes whose padded field size (3 components) or element
// point index (125 per element) would overflow 32-bit device indices.
void check_index_range_3d(int64_t nglob_pad, int64_t nspec);
HostAssembly build_assembly_3d(const sfx_mesh3d &mesh, int nsrc, const sfx_source *src, int nrec,
                               const sfx_receiver *rec, const sfx_run_desc &run);
HostAssembly build_assembly_3d_db(const sfx_mesh3d_db &mesh, int nsrc, const sfx_source *src, int nrec,
                                  const sfx_receiver *rec, const sfx_run_desc &run);

}  // namespace sfx
