// Device kernels of the Newmark step (gfx950, wave64).
//
// One global-point field is stored component-major: comp c of point g is
// f[c * nglob_pad + g].  Element arrays are element-major ([e][p]).
#pragma once

#include <hip/hip_runtime.h>

namespace sfx {

// GLL constants passed by value: they live in scalar registers, so the
// derivative contractions are VALU FMAs with an SGPR operand.
struct GllConst {
  float hp[5][5];   // hprime[i][l] = l_l'(xi_i)
  float hpw[5][5];  // hprime_wgll[i][l] = hprime[l][i] * w[l]  (compute_mesh.hpp:266-272)
  float w[5];
};

// Derivative matrices of a regular hexahedron with its constant geometry
// folded in (built on the host by scaled_gll3): hx = xix hprime, hy = ety
// hprime, hz = gaz hprime for the gradient, wx = J xix hprime_wgll, wy, wz for
// the divergence.  The force kernels then form the strain from the
// contractions directly and store the stress unscaled.
struct Gll3 {
  float hx[5][5], hy[5][5], hz[5][5];
  float wx[5][5], wy[5][5], wz[5][5];
  float w[5];
};
Gll3 scaled_gll3(const GllConst &q, float xix, float ety, float gaz, float jac);

// x + s * y with the product rounded before the sum (never one fused
// multiply-add, whatever contraction the caller's code allows).
__device__ __forceinline__ float plus_product_rounded(float x, float s, float y) {
#pragma clang fp contract(off)
  return x + s * y;
}

struct Fields {
  float *u, *v, *a;
  const float *minv;
  int nglob_pad;
};

// Newmark constants (newmark.hpp:45-46): dt, dt/2, dt^2/2 in type_real.
struct NewmarkConst {
  float dt, dto2, dt2o2;
};

// Stiffness over device element slots [first, first + count).  atomic = false
// requires that no two of these elements share a global point (one color).
// rec_mode 0: plain; 1: store the local forces of Stacey elements
// (rec_index[slot] >= 0) into rec (nstacey, 25, 2); 2: add the stored values
// for those elements instead of their computed forces.
void launch_stiffness_2d(hipStream_t s, int first, int count, bool atomic, const int *gather, const float *xix,
                         const float *xiz, const float *gammax, const float *gammaz, const float *jac,
                         const float *kappa, const float *mu, const GllConst &q, Fields f, int rec_mode = 0,
                         const int *rec_index = nullptr, float *rec = nullptr, const int *step = nullptr,
                         long long rec_stride = 0, const float *geo = nullptr);
// Plain 2-D stiffness with a chosen number of elements per workgroup (2, 5 or
// 10: 64, 128 or 256 threads); the batch autotuner picks it.
void launch_stiffness_2d_tuned(hipStream_t s, int elems_per_block, int first, int count, bool atomic,
                               const int *gather, const float *xix, const float *xiz, const float *gammax,
                               const float *gammaz, const float *jac, const float *kappa, const float *mu,
                               const GllConst &q, Fields f);
// Combined adjoint step: both fields' forces in one pass per color (backward
// replays the recorded Stacey elements; nothing added to it when bwd_active is
// 0); with frechet set, Kmu / Kkappa (nspec, 25) accumulate the strain kernels.
void launch_stiffness_pair_2d(hipStream_t s, int first, int count, bool atomic, const int *gather,
                              const float *xix, const float *xiz, const float *gammax, const float *gammaz,
                              const float *jac, const float *kappa, const float *mu, const GllConst &q, Fields adj,
                              Fields bwd, int bwd_active, const int *rec_index, const float *rec, int frechet,
                              float dt, int nspec, float *Kmu, float *Kkappa, const int *step = nullptr,
                              long long rec_stride = 0);
// *step += delta (the device step counter of graph-captured steps).
void launch_step_tick(hipStream_t s, int *step, int delta);
// Division + corrector of both fields, R += dt a+ . u_b, predictors of the
// next step (+dt adjoint, -dt backward).
void launch_update_pair(hipStream_t s, int ncomp, Fields adj, Fields bwd, NewmarkConst nc, int bwd_correct, float *R);
void launch_rho_product(hipStream_t s, int ncomp, Fields adj, Fields bwd, float dt, float *R);
// out (6, nspec, 25) from R (per point), Kmu, Kkappa.
void launch_frechet_finish(hipStream_t s, int npts, const int *gather, const float *kappa, const float *mu,
                           const float *rho, const float *R, const float *Kmu, const float *Kkappa, float *out);
// K (6, nspec, 25) += Fréchet derivatives of one step from the adjoint and
// backward fields.
void launch_frechet_2d(hipStream_t s, int nspec, const int *gather, const float *xix, const float *xiz,
                       const float *gammax, const float *gammaz, const float *kappa, const float *mu, const float *rho,
                       const GllConst &q, Fields adj, Fields bwd, float dt, float *K);
void launch_stiffness_3d(hipStream_t s, int first, int count, bool atomic, const int *gather, const float *kappa,
                         const float *mu, float xix, float ety, float gaz, float jac, const GllConst &q, Fields f);
// ---- 3-D structured path with buffered element forces (SFX_SCATTER_GATHER) ----
// Tiled lattice of GLL points: point (I, J, K) lives in tile
// (I / 4, J / 4, K / 4) of TX = nx + 1 by TY by TZ tiles, at device index
// ((tz TY + ty) TX + tx) 64 + ((K % 4) 4 + J % 4) 4 + I % 4.
// Element forces re[e * 384 + c * 128 + slot] with the slot order of
// re_slot(): the element's own tile (local i, j, k < 4) first, then its faces,
// edges and corner at local index 4.
struct Lattice {
  int nx, ny, nz, TX, TY, TZ;
  int nint_pad;  // block layout only: slots of the block-interior points (lattice.hpp)
};
// Element forces -K_e u_e of every element into re (plain coalesced stores);
// kappa and mu per point, or (kmu given) one (kappa, mu) pair per element.
void launch_elemforce_3d(hipStream_t s, Lattice lat, const float *kappa, const float *mu, const float *kmu,
                         float xix, float ety, float gaz, float jac, const GllConst &q, const float *u,
                         int nglob_pad, float *re);
// re += source contributions (stf(istep) * source_array) of each source element.
void launch_sources_re(hipStream_t s, int nsrc, int istep, const int *ispec, const float *array, const float *stf,
                       float *re);
// re += Stacey traction of every boundary entry (slot = element * 125 + point);
// entries [seg[i], seg[i+1]) share one slot and are summed by one thread.
void launch_stacey_re(hipStream_t s, int nseg, const int *seg, const int *gather, const int *slot,
                      const float *normal, const float *weight, const float *rho_vp, const float *rho_vs, Fields f,
                      float *re);
// Both of the above in one launch (one extra block for the sources); only
// valid when no source element has an absorbing face.
void launch_stacey_sources_re(hipStream_t s, int nseg, const int *seg, const int *gather, const int *slot,
                              const float *normal, const float *weight, const float *rho_vp, const float *rho_vs,
                              Fields f, int nsrc, int istep, const int *src_ispec, const float *src_array,
                              const float *stf, float *re);
// The same plus one block per receiver for displacement traces (they read only
// u, so they may run beside the Stacey and source blocks).
void launch_stacey_sources_seismo_re(hipStream_t s, int nseg, const int *seg, const int *gather, const int *slot,
                                     const float *normal, const float *weight, const float *rho_vp,
                                     const float *rho_vs, Fields f, int nsrc, int istep, const int *src_ispec,
                                     const float *src_array, const float *stf, float *re, Lattice lat, int nrec,
                                     int ntypes, const int *types, int isig, const int *rec_ispec, const float *rec_h,
                                     float dto2, float *seismo);
// Per lattice point: a = M^-1 sum_e re(e, p); corrector(n) + predictor(n+1) on u, v.
// zero: an all-zero block of 384 floats read in place of absent contributions.
// minv1: M^-1 of one component (equal for all three off the absorbing faces);
// amask: bit f set when face f (x-, x+, y-, y+, z-, z+) is absorbing.
void launch_assemble_update_3d(hipStream_t s, Lattice lat, const float *re, const float *zero, Fields f,
                               NewmarkConst nc, const float *minv1, int amask);
// a = sum_e re(e, p) (times M^-1 when divide) for the seam path and read-out.
void launch_assemble_3d(hipStream_t s, Lattice lat, const float *re, Fields f, int divide);
// Seismograms sampled before the fused update: a from re, v corrected
// (need_a: some trace is a velocity or an acceleration).
void launch_seismograms_re(hipStream_t s, Lattice lat, int nrec, int ntypes, const int *types, int isig,
                           const int *ispec, const float *h, const float *re, Fields f, float dto2, float *seismo,
                           bool need_a);

// ---- 3-D structured path with block assembly (SFX_SCATTER_BLOCK) ----
// Workgroup = 2 x 2 x 2 elements (nx, ny, nz even).  Block-interior points are
// assembled and updated inside launch_block_step_3d; block-surface sums go to
// bs (block_surface_floats floats) and are finished by launch_surface_update_3d.
// store_a: 0 keeps f.a untouched, 1 stores a = M^-1 f in the interior of blocks
// with rec_block set, 2 stores it everywhere.  divide_only: only a = M^-1 f is
// stored (the update_wavefields seam), u and v are left as they are.
size_t block3d_surface_floats(const Lattice &lat);
void launch_block_step_3d(hipStream_t s, Lattice lat, const float *kappa, const float *mu, float xix, float ety,
                          float gaz, float jac, const GllConst &q, Fields f, NewmarkConst nc, int nsrc, int istep,
                          const int *src_ispec, const float *src_array, const float *stf, const int *rec_block,
                          int store_a, int divide_only, float *bs);
void launch_surface_update_3d(hipStream_t s, Lattice lat, const float *bs, Fields f, NewmarkConst nc, int store_a,
                              int divide_only);
// bs += Stacey traction; bslot = owner * 386 + slot (lattice.hpp), entries
// [seg[i], seg[i+1]) share one slot.
void launch_stacey_bs(hipStream_t s, int nseg, const int *seg, const int *gather, const int *bslot,
                      const float *normal, const float *weight, const float *rho_vp, const float *rho_vs, Fields f,
                      float *bs);
// u, v at the receiver elements' points before the block kernel overwrites them.
void launch_receiver_snapshot(hipStream_t s, Lattice lat, int nrec, const int *ispec, Fields f, float *snap);
void launch_seismograms_block(hipStream_t s, Lattice lat, int nrec, int ntypes, const int *types, int isig,
                              const int *ispec, const float *h, const float *bs, const float *snap, Fields f,
                              float dto2, float *seismo);

// rec != nullptr: also add each entry's traction to rec[bnd_rec[entry] * dim + d].
// Kernels given a device step counter `step` read the step from it (and, for
// records, offset rec by step * rec_stride) instead of the explicit argument,
// so that one captured graph serves every step.
void launch_stacey(hipStream_t s, int dim, int nbnd, const int *gather, const float *normal, const float *weight,
                   const float *rho_vp, const float *rho_vs, Fields f, const int *bnd_rec = nullptr,
                   float *rec = nullptr, const int *step = nullptr, long long rec_stride = 0);
// norder >= 0: only the first norder sources of `order` (nsrc stays the
// stride of stf); -1: all nsrc.
void launch_sources(hipStream_t s, int dim, int npe, int nsrc, int istep, const int *order, const int *ispec,
                    const int *gather, const float *array, const float *stf, Fields f, const int *step = nullptr,
                    int norder = -1);
// Per-point 2-D derivatives and properties read by the pressure seismogram.
struct PointGeom2D {
  const float *xix, *xiz, *gammax, *gammaz, *kappa, *mu;
};
// corrected = 1: sample the state after the (not yet applied) mass division and
// corrector of this step, i.e. u, v + dt/2 a M^-1, a M^-1.  Types 0..2
// displacement, velocity, acceleration; 3 pressure (2-D).
// 2-D gather path: element forces re2 (RE2_ELEM floats per element, point p
// component c at 2 p + c) and, per field point g, the re2 offsets of its
// element contributions in [off[g], off[g + 1]) of ent (summed in that order).
constexpr int RE2_ELEM = 52;
struct ForceSum2D {
  const int *off = nullptr, *ent = nullptr;
  const float *re2 = nullptr;
};

void launch_seismograms(hipStream_t s, int dim, int npe, int nrec, int ntypes, const int *types, int isig,
                        const int *ispec, const int *gather, const float *h, Fields f, float dto2, int corrected,
                        float *seismo, const PointGeom2D &pg, const GllConst &q, const int *step = nullptr,
                        int nsb = 1, int sig_top = -1, ForceSum2D fs = ForceSum2D{}, const int *rec_list = nullptr,
                        int nlist = -1);
// Where elemforce_2d takes an element point's geometry from.  Every form gives
// the same bits: the control-node forms repeat the host's float operations in
// the host's order (compute_partial_derivatives.cpp:46-91, no fused
// multiply-add, correctly rounded division), so they trade 20 bytes per point
// for 8 or 18 floats per element.
enum Geom2DMode : int { GEOM2D_POINT = 0, GEOM2D_RECORD = 1, GEOM2D_CTRL4 = 4, GEOM2D_CTRL9 = 9 };
struct Geom2D {
  int mode = GEOM2D_POINT;
  const float *xix = nullptr, *xiz = nullptr, *gammax = nullptr, *gammaz = nullptr, *jac = nullptr;  // per point
  const float *record = nullptr;  // per element: xix, xiz, gammax, gammaz, J, 3 unused
  const float *ctrl = nullptr;    // per element: its ngnod control nodes (x, z)
  const float *dsh = nullptr;     // (25, 2, ngnod): shape derivatives in xi, then gamma, at each GLL point
};
// kappa and mu per local point, or (kmu given) one (kappa, mu) pair per element
void launch_elemforce_2d(hipStream_t s, int nspec, const int *gather, const Geom2D &gm, const float *kappa,
                         const float *mu, const float *kmu, const GllConst &q, const float *u, int nglob_pad,
                         float *re2);
void launch_stacey_re2(hipStream_t s, int nseg, const int *seg, const int *gather, const int *slot,
                       const float *normal, const float *weight, const float *rho_vp, const float *rho_vs, Fields f,
                       float *re2);
void launch_sources_re2(hipStream_t s, int nsrc, int istep, const int *order, const int *ispec, const float *array,
                        const float *stf, float *re2, const int *step = nullptr);
void launch_assemble_update_2d(hipStream_t s, int nglob, ForceSum2D fs, Fields f, NewmarkConst nc, int divide_only);
void launch_update_fused(hipStream_t s, int ncomp, Fields f, NewmarkConst nc);
// rounded: every product and sum rounded on its own (no fused multiply-add),
// as the 2-D gather path's fused update rounds them; its seams pass true.
void launch_predictor(hipStream_t s, int ncomp, Fields f, NewmarkConst nc, bool rounded = false);
void launch_divide_corrector(hipStream_t s, int ncomp, Fields f, NewmarkConst nc, int divide, int correct,
                             bool rounded = false);
void launch_fill(hipStream_t s, float *p, size_t n, float value);

// ---- 3-D general hexahedra (general3d.hip) ----
// Stiffness over device element slots [first, first + count) with per-point
// geometry geo (nspec, 10, 125): xix, xiy, xiz, etax, etay, etaz, gammax,
// gammay, gammaz, jacobian.
void launch_stiffness_3d_general(hipStream_t s, int first, int count, bool atomic, const int *gather,
                                 const float *geo, const float *kappa, const float *mu, const GllConst &q,
                                 Fields f);

// ---- 2-D acoustic medium and elastic-acoustic coupling (acoustic.hip) ----
// Acoustic stiffness over device element slots [first, first + count) of the
// acoustic field f: rho^-1 grad chi, the Stacey term stacey[p] * chi_dot
// (w |n| rho_vpinverse per point, 0 off absorbing edges) and zero at
// free-surface points (dirichlet[p] = 1).
void launch_acoustic_stiffness_2d(hipStream_t s, int first, int count, bool atomic, const int *gather,
                                  const float *xix, const float *xiz, const float *gammax, const float *gammaz,
                                  const float *jac, const float *rho_inv, const float *stacey, const int *dirichlet,
                                  const GllConst &q, Fields f);
// Interface terms: to_elastic = 0 adds factor (nx u_x + nz u_z) of src (the
// elastic field) to dst (the acoustic field); to_elastic = 1 adds factor n
// chi_dot_dot of src (acoustic) to dst (elastic).  coef (n, 3): factor, nx, nz.
void launch_coupling_2d(hipStream_t s, int n, const int *self, const int *other, const float *coef, Fields dst,
                        Fields src, int to_elastic);
// Acoustic sources order[0..norder): a += stf * array / kappa.
void launch_acoustic_sources(hipStream_t s, int norder, int nsrc, int istep, const int *order, const int *ispec,
                             const int *gather, const float *array, const float *stf, const float *kappa, Fields f);
// Seismograms of the receivers list[0..nlist) in acoustic elements.
void launch_acoustic_seismograms(hipStream_t s, int nlist, const int *list, int nrec, int ntypes, const int *types,
                                 int isig, const int *ispec, const int *gather, const float *h, const float *xix,
                                 const float *xiz, const float *gammax, const float *gammaz, const float *rho_inv,
                                 const GllConst &q, Fields f, float *seismo);

}  // namespace sfx
