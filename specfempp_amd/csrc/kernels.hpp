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

struct Fields {
  float *u, *v, *a;
  const float *minv;
  int nglob_pad;
};

// Newmark constants (newmark.hpp:45-46): dt, dt/2, dt^2/2 in type_real.
struct NewmarkConst {
  float dt, dto2, dt2o2;
};

void launch_stiffness_2d(hipStream_t s, int nspec, const int *gather, const float *xix, const float *xiz,
                         const float *gammax, const float *gammaz, const float *jac, const float *kappa,
                         const float *mu, const GllConst &q, Fields f);
void launch_stiffness_3d(hipStream_t s, int nspec, const int *gather, const float *kappa, const float *mu,
                         float xix, float ety, float gaz, float jac, const GllConst &q, Fields f);
void launch_stacey(hipStream_t s, int dim, int nbnd, const int *gather, const float *normal, const float *weight,
                   const float *rho_vp, const float *rho_vs, Fields f);
void launch_sources(hipStream_t s, int dim, int npe, int nsrc, int istep, const int *order, const int *ispec,
                    const int *gather, const float *array, const float *stf, Fields f);
// corrected = 1: sample the state after the (not yet applied) mass division and
// corrector of this step, i.e. u, v + dt/2 a M^-1, a M^-1.
void launch_seismograms(hipStream_t s, int dim, int npe, int nrec, int ntypes, const int *types, int isig,
                        const int *ispec, const int *gather, const float *h, Fields f, float dto2, int corrected,
                        float *seismo);
void launch_update_fused(hipStream_t s, int ncomp, Fields f, NewmarkConst nc);
void launch_predictor(hipStream_t s, int ncomp, Fields f, NewmarkConst nc);
void launch_divide_corrector(hipStream_t s, int ncomp, Fields f, NewmarkConst nc, int divide, int correct);
void launch_fill(hipStream_t s, float *p, size_t n, float value);

}  // namespace sfx
