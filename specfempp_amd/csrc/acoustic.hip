// 2-D acoustic medium and elastic-acoustic coupling kernels (gfx950, wave64).
//
// Reference kernels these replace:
//   acoustic stiffness   compute_stiffness_interaction.tpp:143-259 with the
//                        acoustic stress rho^-1 grad chi
//                        (include/medium/dim2/acoustic/isotropic/stress.hpp),
//                        the acoustic Stacey term (stacey.tpp:172-213) and the
//                        free surface (boundary_conditions/dirichlet: the
//                        point's contribution is set to zero)
//   acoustic sources     compute_source_interaction.tpp:80-95 with
//                        include/medium/dim2/acoustic/isotropic/source.hpp
//                        (stf * lagrange / kappa)
//   coupling             include/coupled_interface/coupled_interface.tpp and
//                        interface_kernels.hpp:18-54 (acoustic side: factor
//                        n . u; elastic side: factor n chi_dot_dot)
//   seismograms          compute_seismogram.tpp:103-135 with
//                        include/medium/dim2/acoustic/isotropic/wavefield.hpp
//                        (grad chi / rho; pressure -chi_dot_dot)
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace sfx {

namespace {

constexpr int EA = 5;  // acoustic elements per 128-thread workgroup

inline unsigned blocks_of(long long n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

template <bool ATOMIC> __device__ __forceinline__ void add_to(float *p, float v) {
  if constexpr (ATOMIC)
    unsafeAtomicAdd(p, v);
  else
    *p += v;
}

// One thread per GLL point of EA elements.  chi and the integrand F are
// staged in LDS; the derivative matrices' rows are held in registers.
template <bool ATOMIC>
__global__ __launch_bounds__(128) void acoustic_stiffness_2d_kernel(
    int first, int count, const int *__restrict__ gather, const float *__restrict__ xix_,
    const float *__restrict__ xiz_, const float *__restrict__ gammax_, const float *__restrict__ gammaz_,
    const float *__restrict__ jac_, const float *__restrict__ rho_inv_, const float *__restrict__ stacey_,
    const int *__restrict__ dirichlet_, GllConst q, Fields f) {
  __shared__ float sc[EA][25];
  __shared__ float sF[EA][2][25];
  const int t = threadIdx.x;
  const int el = t / 25;
  const int p = t - el * 25;
  const int iz = p / 5, ix = p - (p / 5) * 5;
  const int el_idx = blockIdx.x * EA + el;
  const bool active = (el < EA) && (el_idx < count);
  const size_t idx = static_cast<size_t>(first + el_idx) * 25 + p;
  int g = 0;
  if (active) {
    g = gather[idx];
    sc[el][p] = f.u[g];
  }
  __syncthreads();
  if (active) {
    float dxi = 0.0f, dga = 0.0f;
#pragma unroll
    for (int l = 0; l < 5; l++) {
      dxi += q.hp[ix][l] * sc[el][iz * 5 + l];
      dga += q.hp[iz][l] * sc[el][l * 5 + ix];
    }
    const float xix = xix_[idx], xiz = xiz_[idx], gammax = gammax_[idx], gammaz = gammaz_[idx], J = jac_[idx];
    const float ri = rho_inv_[idx];
    const float du0 = xix * dxi + gammax * dga;
    const float du1 = xiz * dxi + gammaz * dga;
    const float T0 = ri * du0, T1 = ri * du1;
    // core/specfem/point/stress.hpp:118-139 for one component
    sF[el][0][p] = J * (T0 * xix + T1 * xiz);
    sF[el][1][p] = J * (T0 * gammax + T1 * gammaz);
  }
  __syncthreads();
  if (!active) return;
  // include/algorithms/divergence.hpp:86-98
  float t1 = 0.0f, t2 = 0.0f;
#pragma unroll
  for (int l = 0; l < 5; l++) {
    t1 += sF[el][0][iz * 5 + l] * q.hpw[ix][l];
    t2 += sF[el][1][l * 5 + ix] * q.hpw[iz][l];
  }
  float r = -(q.w[iz] * t1 + q.w[ix] * t2);
  const float c = stacey_[idx];
  if (c != 0.0f) r += -c * f.v[g];
  if (dirichlet_[idx]) r = 0.0f;
  add_to<ATOMIC>(&f.a[g], r);
}

// a[self] += factor (nx u_x[other] + nz u_z[other]) into the acoustic field
// from the elastic displacement (to_elastic = 0), or
// a_x[self] += factor nx chi_dot_dot[other], a_z[self] += factor nz
// chi_dot_dot[other] into the elastic field (to_elastic = 1).
__global__ __launch_bounds__(256) void coupling_2d_kernel(int n, const int *__restrict__ self,
                                                          const int *__restrict__ other,
                                                          const float *__restrict__ coef, Fields dst, Fields src,
                                                          int to_elastic) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = self[i], o = other[i];
  const float factor = coef[3 * i], nx = coef[3 * i + 1], nz = coef[3 * i + 2];
  if (to_elastic) {
    const float chi_dd = src.a[o];
    unsafeAtomicAdd(&dst.a[s], factor * nx * chi_dd);
    unsafeAtomicAdd(&dst.a[dst.nglob_pad + s], factor * nz * chi_dd);
  } else {
    unsafeAtomicAdd(&dst.a[s], factor * (nx * src.u[o] + nz * src.u[src.nglob_pad + o]));
  }
}

// One workgroup per acoustic source (in application order), one thread per
// element point.
__global__ __launch_bounds__(64) void acoustic_sources_kernel(int norder, int nsrc, int istep,
                                                              const int *__restrict__ order,
                                                              const int *__restrict__ ispec,
                                                              const int *__restrict__ gather,
                                                              const float *__restrict__ array,
                                                              const float *__restrict__ stf,
                                                              const float *__restrict__ kappa, Fields f) {
  const int p = threadIdx.x, k = blockIdx.x;
  if (p >= 25 || k >= norder) return;
  const int s = order[k];
  const float sv = stf[static_cast<size_t>(istep) * nsrc + s];
  const size_t idx = static_cast<size_t>(ispec[s]) * 25 + p;
  unsafeAtomicAdd(&f.a[gather[idx]], sv * array[static_cast<size_t>(s) * 50 + p] / kappa[idx]);
}

// One workgroup per receiver of `list`: the element's potential (or its time
// derivatives) in LDS, the gradient times 1/rho at every point, Lagrange
// weights, wave reduction.  Pressure is -chi_dot_dot.
__global__ __launch_bounds__(64) void acoustic_seismogram_kernel(
    const int *__restrict__ list, int nrec, int ntypes, const int *__restrict__ types, int isig,
    const int *__restrict__ ispec, const int *__restrict__ gather, const float *__restrict__ h,
    const float *__restrict__ xix_, const float *__restrict__ xiz_, const float *__restrict__ gammax_,
    const float *__restrict__ gammaz_, const float *__restrict__ rho_inv_, GllConst q, Fields f,
    float *__restrict__ seismo) {
  __shared__ float sc[25];
  const int r = list[blockIdx.x];
  const int p = threadIdx.x;
  const int iz = p / 5, ix = p - (p / 5) * 5;
  const size_t idx = static_cast<size_t>(ispec[r]) * 25 + (p < 25 ? p : 0);
  const int g = p < 25 ? gather[idx] : 0;
  const float hw = p < 25 ? h[static_cast<size_t>(r) * 25 + p] : 0.0f;
  for (int it = 0; it < ntypes; it++) {
    const int type = types[it];
    const float *fld = type == 0 ? f.u : (type == 1 ? f.v : f.a);
    __syncthreads();
    if (p < 25) sc[p] = fld[g];
    __syncthreads();
    float v0 = 0.0f, v1 = 0.0f;
    if (p < 25) {
      if (type == 3) {
        v0 = -sc[p];
      } else {
        float dxi = 0.0f, dga = 0.0f;
#pragma unroll
        for (int l = 0; l < 5; l++) {
          dxi += q.hp[ix][l] * sc[iz * 5 + l];
          dga += q.hp[iz][l] * sc[l * 5 + ix];
        }
        const float ri = rho_inv_[idx];
        v0 = ri * (xix_[idx] * dxi + gammax_[idx] * dga);
        v1 = ri * (xiz_[idx] * dxi + gammaz_[idx] * dga);
      }
      v0 *= hw;
      v1 *= hw;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v0 += __shfl_down(v0, off, 64);
      v1 += __shfl_down(v1, off, 64);
    }
    if (p == 0) {
      float *o = seismo + ((static_cast<size_t>(isig) * ntypes + it) * nrec + r) * 2;
      o[0] = v0;
      o[1] = type == 3 ? 0.0f : v1;
    }
  }
}

}  // namespace

void launch_acoustic_stiffness_2d(hipStream_t s, int first, int count, bool atomic, const int *gather,
                                  const float *xix, const float *xiz, const float *gammax, const float *gammaz,
                                  const float *jac, const float *rho_inv, const float *stacey, const int *dirichlet,
                                  const GllConst &q, Fields f) {
  if (count <= 0) return;
  if (atomic)
    hipLaunchKernelGGL(acoustic_stiffness_2d_kernel<true>, dim3(blocks_of(count, EA)), dim3(128), 0, s, first, count,
                       gather, xix, xiz, gammax, gammaz, jac, rho_inv, stacey, dirichlet, q, f);
  else
    hipLaunchKernelGGL(acoustic_stiffness_2d_kernel<false>, dim3(blocks_of(count, EA)), dim3(128), 0, s, first,
                       count, gather, xix, xiz, gammax, gammaz, jac, rho_inv, stacey, dirichlet, q, f);
}

void launch_coupling_2d(hipStream_t s, int n, const int *self, const int *other, const float *coef, Fields dst,
                        Fields src, int to_elastic) {
  if (n <= 0) return;
  hipLaunchKernelGGL(coupling_2d_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, s, n, self, other, coef, dst, src,
                     to_elastic);
}

void launch_acoustic_sources(hipStream_t s, int norder, int nsrc, int istep, const int *order, const int *ispec,
                             const int *gather, const float *array, const float *stf, const float *kappa, Fields f) {
  if (norder <= 0) return;
  hipLaunchKernelGGL(acoustic_sources_kernel, dim3(norder), dim3(64), 0, s, norder, nsrc, istep, order, ispec, gather,
                     array, stf, kappa, f);
}

void launch_acoustic_seismograms(hipStream_t s, int nlist, const int *list, int nrec, int ntypes, const int *types,
                                 int isig, const int *ispec, const int *gather, const float *h, const float *xix,
                                 const float *xiz, const float *gammax, const float *gammaz, const float *rho_inv,
                                 const GllConst &q, Fields f, float *seismo) {
  if (nlist <= 0 || ntypes <= 0) return;
  hipLaunchKernelGGL(acoustic_seismogram_kernel, dim3(nlist), dim3(64), 0, s, list, nrec, ntypes, types, isig, ispec,
                     gather, h, xix, xiz, gammax, gammaz, rho_inv, q, f, seismo);
}

}  // namespace sfx
