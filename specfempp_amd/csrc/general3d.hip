// 3-D stiffness on general hexahedra (mesh databases with irregular
// elements): the weak-form internal force of
// compute_stiffness_interaction.tpp:143-259 extended to (xi, eta, gamma)
// with the full per-point derivative matrix (xix ... gammaz) and jacobian of
// the 3-D database (src/io/mesh/impl/fortran/dim3/mesh.cpp:208-265):
//   du_c/dx_r = sum_s (du_c/dxi_s) dxi_s/dx_r           (gradient.hpp:81-103)
//   sigma = lambda tr(eps) I + 2 mu eps                  (isotropic, kappa + 4/3 mu)
//   F_cs = J sum_r sigma_cr dxi_s/dx_r                   (point/stress.hpp:118-139)
//   r_c = -sum_s w w sum_l F_cs(l) hprime_wgll           (divergence.hpp:86-98)
// One thread owns an (i, j) column of an element and its 5 points along k;
// the xi and eta integrands go through LDS, the gamma ones stay in registers.
// 5 elements per 128-thread workgroup.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace sfx {

namespace {

constexpr int EG = 5;

template <bool ATOMIC> __device__ __forceinline__ void add_force(float *p, float v) {
  if constexpr (ATOMIC)
    unsafeAtomicAdd(p, v);
  else
    *p += v;
}

inline unsigned groups(long long n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

template <bool ATOMIC>
__global__ __launch_bounds__(128) void stiffness_3d_general_kernel(int first, int count,
                                                                   const int *__restrict__ gather,
                                                                   const float *__restrict__ geo,
                                                                   const float *__restrict__ kappa_,
                                                                   const float *__restrict__ mu_, GllConst q,
                                                                   Fields f) {
  __shared__ float su[EG][3][125];
  __shared__ float sF0[EG][3][125];  // F(c, xi)
  __shared__ float sF1[EG][3][125];  // F(c, eta)

  const int t = threadIdx.x;
  const int el = t / 25;
  const int col = t - el * 25;
  const int j = col / 5, i = col - (col / 5) * 5;
  const int el_idx = blockIdx.x * EG + el;
  const bool active = (el < EG) && (el_idx < count);
  const int e = first + el_idx;
  const size_t base = static_cast<size_t>(e) * 125 + col;
  const float *ge = geo + static_cast<size_t>(e) * 1250 + col;
  const int np = f.nglob_pad;

  int g[5];
  float u[5][3];
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++) g[k] = gather[base + k * 25];
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        u[k][c] = f.u[c * np + g[k]];
        su[el][c][k * 25 + col] = u[k][c];
      }
  }
  __syncthreads();

  float F2[5][3];  // F(c, gamma), in registers
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      float dr[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // dr[c][s] = du_c / dxi_s
#pragma unroll
      for (int l = 0; l < 5; l++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          dr[c][0] += q.hp[i][l] * su[el][c][k * 25 + j * 5 + l];
          dr[c][1] += q.hp[j][l] * su[el][c][k * 25 + l * 5 + i];
          dr[c][2] += q.hp[k][l] * u[l][c];
        }
      }
      float d9[9];
#pragma unroll
      for (int d = 0; d < 9; d++) d9[d] = ge[d * 125 + k * 25];
      const float J = ge[9 * 125 + k * 25];
      float du[3][3];  // du[c][r] = du_c / dx_r
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) du[c][r] = dr[c][0] * d9[0 + r] + dr[c][1] * d9[3 + r] + dr[c][2] * d9[6 + r];
      const float kappa = kappa_[base + k * 25];
      const float mu = mu_[base + k * 25];
      const float lp2m = kappa + (4.0f / 3.0f) * mu;
      const float lambda = lp2m - 2.0f * mu;
      const float tr = du[0][0] + du[1][1] + du[2][2];
      float s[3][3];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) s[c][r] = mu * (du[c][r] + du[r][c]) + (c == r ? lambda * tr : 0.0f);
      const int p = k * 25 + col;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        sF0[el][c][p] = J * (s[c][0] * d9[0] + s[c][1] * d9[1] + s[c][2] * d9[2]);
        sF1[el][c][p] = J * (s[c][0] * d9[3] + s[c][1] * d9[4] + s[c][2] * d9[5]);
        F2[k][c] = J * (s[c][0] * d9[6] + s[c][1] * d9[7] + s[c][2] * d9[8]);
      }
    }
  }
  __syncthreads();
  if (!active) return;
  const float wi = q.w[i], wj = q.w[j];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    float r0[3] = {0, 0, 0}, r1[3] = {0, 0, 0}, r2[3] = {0, 0, 0};
#pragma unroll
    for (int l = 0; l < 5; l++) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        r0[c] += sF0[el][c][k * 25 + j * 5 + l] * q.hpw[i][l];
        r1[c] += sF1[el][c][k * 25 + l * 5 + i] * q.hpw[j][l];
        r2[c] += F2[l][c] * q.hpw[k][l];
      }
    }
    const float wjk = wj * q.w[k], wik = wi * q.w[k], wij = wi * wj;
#pragma unroll
    for (int c = 0; c < 3; c++) add_force<ATOMIC>(&f.a[c * np + g[k]], -(wjk * r0[c] + wik * r1[c] + wij * r2[c]));
  }
}

}  // namespace

void launch_stiffness_3d_general(hipStream_t s, int first, int count, bool atomic, const int *gather,
                                 const float *geo, const float *kappa, const float *mu, const GllConst &q,
                                 Fields f) {
  if (count <= 0) return;
  if (atomic)
    hipLaunchKernelGGL(stiffness_3d_general_kernel<true>, dim3(groups(count, EG)), dim3(128), 0, s, first, count,
                       gather, geo, kappa, mu, q, f);
  else
    hipLaunchKernelGGL(stiffness_3d_general_kernel<false>, dim3(groups(count, EG)), dim3(128), 0, s, first, count,
                       gather, geo, kappa, mu, q, f);
}

}  // namespace sfx
