// Newmark-step kernels for MI355X (gfx950, wave64).
//
// Reference kernels these replace (include/kokkos_kernels/impl/):
//   stiffness_2d / stiffness_3d  compute_stiffness_interaction.tpp:143-259
//   stacey                       the Stacey branch of the same kernel
//                                (boundary_conditions/stacey/stacey.tpp:36-72)
//   sources                      compute_source_interaction.tpp:80-95
//   seismograms                  compute_seismogram.tpp:103-135
//   update_fused                 divide_mass_matrix.tpp:42-49 + newmark.tpp:11-116
//                                (corrector of step n and predictor of n+1 in one pass)
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "kernels.hpp"
#include "lattice.hpp"

namespace sfx {

namespace {

__device__ __forceinline__ void atomic_add(float *p, float v) { unsafeAtomicAdd(p, v); }

// Element-force scatter: a float atomic, or a plain read-modify-write when the
// launch covers one color (no other thread of the launch touches the point).
template <bool ATOMIC> __device__ __forceinline__ void scatter_add(float *p, float v) {
  if constexpr (ATOMIC)
    unsafeAtomicAdd(p, v);
  else
    *p += v;
}

// ---------------------------------------------------------------------------
// 2-D stiffness: 5 elements per 128-thread block, one thread per GLL point.
// The element's displacement and its stress integrand F are staged in LDS
// (the reference's chunk_element field and stress_integrand).
// ---------------------------------------------------------------------------
constexpr int E2 = 5;  // elements per block

// Stacey-element record (compute_stiffness_interaction.tpp:119-136,251-255):
// rec_mode 1 stores each Stacey element's local force (the Stacey traction is
// added by the boundary kernel), rec_mode 2 replaces it by the stored value
// (the backward field).  rec_index: Stacey element index of each slot, or -1;
// rec: (nstacey, 25, 2) of the current step.
// Threads of a block holding E elements: E * 25 rounded up to whole wavefronts.
constexpr int threads_for(int E) { return (E * 25 + 63) / 64 * 64; }

template <bool ATOMIC, int E = E2>
__global__ __launch_bounds__(threads_for(E)) void stiffness_2d_kernel(
    int first, int count, const int *__restrict__ gather, const float *__restrict__ xix_, const float *__restrict__ xiz_,
    const float *__restrict__ gammax_, const float *__restrict__ gammaz_, const float *__restrict__ jac_,
    const float *__restrict__ kappa_, const float *__restrict__ mu_, GllConst q, Fields f, int rec_mode,
    const int *__restrict__ rec_index, float *__restrict__ rec, const int *__restrict__ step, long long rec_stride,
    const float *__restrict__ geo) {
  __shared__ float su[E][2][25];
  __shared__ float sF[E][4][25];  // F(c, d) as [2c + d]

  const int t = threadIdx.x;
  const int el = t / 25;
  const int p = t - el * 25;
  const int iz = p / 5, ix = p - (p / 5) * 5;
  const int el_idx = blockIdx.x * E + el;
  const bool active = (el < E) && (el_idx < count);
  // Every lane issues its loads, inactive ones re-reading the last element of
  // the range (all addresses valid), and all before the barrier: with no
  // branch around them they are in flight together.
  const int e = first + (el_idx < count ? el_idx : count - 1);
  const size_t idx = static_cast<size_t>(e) * 25 + p;

  const int g = gather[idx];
  const float u0 = f.u[g], u1 = f.u[f.nglob_pad + g];
  // geometry: one record per element when every element's geometry is
  // constant over its points, else the per-point arrays
  const size_t gr = static_cast<size_t>(e) * 8;
  const float xix = geo ? geo[gr] : xix_[idx];
  const float xiz = geo ? geo[gr + 1] : xiz_[idx];
  const float gammax = geo ? geo[gr + 2] : gammax_[idx];
  const float gammaz = geo ? geo[gr + 3] : gammaz_[idx];
  const float J = geo ? geo[gr + 4] : jac_[idx];
  const float kappa = kappa_[idx], mu = mu_[idx];
  // Per-thread rows of the derivative matrices, loaded once (the row index is
  // lane-dependent, so they are held in VGPRs rather than re-read).
  float hpx[5], hpz[5], hwx[5], hwz[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    hpx[l] = q.hp[ix][l];
    hpz[l] = q.hp[iz][l];
    hwx[l] = q.hpw[ix][l];
    hwz[l] = q.hpw[iz][l];
  }
  const float wx = q.w[ix], wz = q.w[iz];
  if (active) {
    su[el][0][p] = u0;
    su[el][1][p] = u1;
  }
  __syncthreads();

  if (active) {
    float dxi[2] = {0.0f, 0.0f}, dga[2] = {0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < 5; l++) {
#pragma unroll
      for (int c = 0; c < 2; c++) {
        dxi[c] += hpx[l] * su[el][c][iz * 5 + l];
        dga[c] += hpz[l] * su[el][c][l * 5 + ix];
      }
    }
    float du[2][2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      du[c][0] = xix * dxi[c] + gammax * dga[c];
      du[c][1] = xiz * dxi[c] + gammaz * dga[c];
    }
    // include/medium/dim2/elastic/isotropic/stress.hpp:25-48 with
    // lambda + 2 mu = kappa + 4/3 mu (core/specfem/point/properties.hpp:141-161)
    const float lp2m = kappa + (4.0f / 3.0f) * mu;
    const float lambda = lp2m - 2.0f * mu;
    const float sxx = lp2m * du[0][0] + lambda * du[1][1];
    const float szz = lp2m * du[1][1] + lambda * du[0][0];
    const float sxz = mu * (du[0][1] + du[1][0]);
    // core/specfem/point/stress.hpp:129-136
    sF[el][0][p] = J * (sxx * xix + sxz * xiz);
    sF[el][1][p] = J * (sxx * gammax + sxz * gammaz);
    sF[el][2][p] = J * (sxz * xix + szz * xiz);
    sF[el][3][p] = J * (sxz * gammax + szz * gammaz);
  }
  __syncthreads();
  if (!active) return;
  // include/algorithms/divergence.hpp:86-98
  float t1[2] = {0.0f, 0.0f}, t2[2] = {0.0f, 0.0f};
#pragma unroll
  for (int l = 0; l < 5; l++) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      t1[c] += sF[el][2 * c][iz * 5 + l] * hwx[l];
      t2[c] += sF[el][2 * c + 1][l * 5 + ix] * hwz[l];
    }
  }
  float r[2];
#pragma unroll
  for (int c = 0; c < 2; c++) r[c] = -(wz * t1[c] + wx * t2[c]);
  if (rec_mode) {
    const int ri = rec_index[e];
    if (ri >= 0) {
      float *rp = rec + (step ? *step * rec_stride : 0) + (static_cast<size_t>(ri) * 25 + p) * 2;
      if (rec_mode == 1) {
        rp[0] = r[0];
        rp[1] = r[1];
      } else {
        r[0] = rp[0];
        r[1] = rp[1];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 2; c++) scatter_add<ATOMIC>(&f.a[c * f.nglob_pad + g], r[c]);
}

// ---------------------------------------------------------------------------
// 2-D gather path (forward runs): the element forces of 5 elements per
// workgroup are staged in LDS and stored with coalesced 16-byte writes
// (RE2_ELEM floats per element); Stacey tractions and sources are added to
// them; one update per field point then sums its contributions in a fixed
// order (ascending element, then point) and applies the Newmark update, with
// no scattered read-modify-write of the acceleration.
// ---------------------------------------------------------------------------
// Geometry of local point p (index idx) of element e from the source GEOM
// names.  The control-node forms sum the shape derivatives times the nodes in
// node order, then form J and the four quotients exactly as the host does.
template <int GEOM>
__device__ __forceinline__ void point_geometry_2d(const Geom2D &gm, int e, int p, size_t idx, float &xix,
                                                  float &xiz, float &gammax, float &gammaz, float &J) {
#pragma clang fp contract(off)
  if constexpr (GEOM == GEOM2D_POINT) {
    xix = gm.xix[idx];
    xiz = gm.xiz[idx];
    gammax = gm.gammax[idx];
    gammaz = gm.gammaz[idx];
    J = gm.jac[idx];
  } else if constexpr (GEOM == GEOM2D_RECORD) {
    const float *ge = gm.record + static_cast<size_t>(e) * 8;
    xix = ge[0];
    xiz = ge[1];
    gammax = ge[2];
    gammaz = ge[3];
    J = ge[4];
  } else {
    constexpr int NG = GEOM;
    const float *c = gm.ctrl + static_cast<size_t>(e) * (2 * NG);
    const float *d = gm.dsh + p * (2 * NG);
    float xxi = 0.0f, zxi = 0.0f, xgamma = 0.0f, zgamma = 0.0f;
#pragma unroll
    for (int n = 0; n < NG; n++) {
      const float cx = c[2 * n], cz = c[2 * n + 1];
      xxi += d[n] * cx;
      zxi += d[n] * cz;
      xgamma += d[NG + n] * cx;
      zgamma += d[NG + n] * cz;
    }
    J = xxi * zgamma - xgamma * zxi;
    xix = zgamma / J;
    gammax = -zxi / J;
    xiz = -xgamma / J;
    gammaz = xxi / J;
  }
}

// A workgroup of 128 threads computes G groups of E2 elements: thread t holds
// local point t % 25 of element t / 25 of every group, and the loads of all
// groups are issued before the first is used, so a wave keeps G times the
// bytes in flight (the kernel is bound by the latency of its two dependent
// loads, the gather index and the displacement it names).
template <int GEOM, bool EPROP, int G>
__global__ __launch_bounds__(128) void elemforce_2d_kernel(int nspec, const int *__restrict__ gather, Geom2D gm,
                                                           const float *__restrict__ kappa_,
                                                           const float *__restrict__ mu_,
                                                           const float *__restrict__ kmu, GllConst q,
                                                           const float *__restrict__ uf, int np,
                                                           float *__restrict__ re2) {
  // every product and sum rounded on its own: the one-launch batch step
  // performs the same operations and must round them the same way
#pragma clang fp contract(off)
  constexpr int NE = G * E2;  // elements per workgroup
  __shared__ float su[NE][2][25];
  __shared__ float sF[NE][4][25];  // F(c, d) as [2c + d]
  __shared__ __align__(16) float st[NE * RE2_ELEM];

  const int t = threadIdx.x;
  const int el = t / 25;
  const int p = t - el * 25;
  const int iz = p / 5, ix = p - (p / 5) * 5;
  const int first = blockIdx.x * NE;
  // Loads are issued for every lane, inactive ones re-reading the last
  // element (all addresses stay in bounds): with no branch around them the
  // loads of all groups are in flight together instead of one after another.
  int e[G];
  bool active[G];
  size_t idx[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    const int eg = first + g * E2 + el;
    active[g] = (el < E2) && (eg < nspec);
    e[g] = eg < nspec ? eg : nspec - 1;
    idx[g] = static_cast<size_t>(e[g]) * 25 + p;
  }

  // the gather indices, then the displacement they name, of every group
  int gi[G];
#pragma unroll
  for (int g = 0; g < G; g++) gi[g] = gather[idx[g]];
  float ux[G], uz[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    ux[g] = uf[gi[g]];
    uz[g] = uf[np + gi[g]];
  }
  float kappa[G], mu[G], xix[G], xiz[G], gammax[G], gammaz[G], J[G];
#pragma unroll
  for (int g = 0; g < G; g++) {
    if constexpr (EPROP) {
      kappa[g] = kmu[2 * static_cast<size_t>(e[g])];
      mu[g] = kmu[2 * static_cast<size_t>(e[g]) + 1];
    } else {
      kappa[g] = kappa_[idx[g]];
      mu[g] = mu_[idx[g]];
    }
    point_geometry_2d<GEOM>(gm, e[g], p, idx[g], xix[g], xiz[g], gammax[g], gammaz[g], J[g]);
  }
  // the derivative rows of this point (lane-indexed, so vector loads; before
  // the barrier, where their latency overlaps the loads above)
  float hpx[5], hpz[5], hwx[5], hwz[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    hpx[l] = q.hp[ix][l];
    hpz[l] = q.hp[iz][l];
    hwx[l] = q.hpw[ix][l];
    hwz[l] = q.hpw[iz][l];
  }
  const float wx = q.w[ix], wz = q.w[iz];
#pragma unroll
  for (int g = 0; g < G; g++)
    if (active[g]) {
      su[g * E2 + el][0][p] = ux[g];
      su[g * E2 + el][1][p] = uz[g];
    }
  __syncthreads();
#pragma unroll
  for (int g = 0; g < G; g++) {
    if (!active[g]) continue;
    const int le = g * E2 + el;
    // the gradient, stress and stress integrand of stiffness_2d_kernel
    float dxi[2] = {0.0f, 0.0f}, dga[2] = {0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < 5; l++) {
#pragma unroll
      for (int c = 0; c < 2; c++) {
        dxi[c] += hpx[l] * su[le][c][iz * 5 + l];
        dga[c] += hpz[l] * su[le][c][l * 5 + ix];
      }
    }
    float du[2][2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      du[c][0] = xix[g] * dxi[c] + gammax[g] * dga[c];
      du[c][1] = xiz[g] * dxi[c] + gammaz[g] * dga[c];
    }
    const float lp2m = kappa[g] + (4.0f / 3.0f) * mu[g];
    const float lambda = lp2m - 2.0f * mu[g];
    const float sxx = lp2m * du[0][0] + lambda * du[1][1];
    const float szz = lp2m * du[1][1] + lambda * du[0][0];
    const float sxz = mu[g] * (du[0][1] + du[1][0]);
    sF[le][0][p] = J[g] * (sxx * xix[g] + sxz * xiz[g]);
    sF[le][1][p] = J[g] * (sxx * gammax[g] + sxz * gammaz[g]);
    sF[le][2][p] = J[g] * (sxz * xix[g] + szz * xiz[g]);
    sF[le][3][p] = J[g] * (sxz * gammax[g] + szz * gammaz[g]);
  }
  __syncthreads();
#pragma unroll
  for (int g = 0; g < G; g++) {
    if (!active[g]) continue;
    const int le = g * E2 + el;
    float t1[2] = {0.0f, 0.0f}, t2[2] = {0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < 5; l++) {
#pragma unroll
      for (int c = 0; c < 2; c++) {
        t1[c] += sF[le][2 * c][iz * 5 + l] * hwx[l];
        t2[c] += sF[le][2 * c + 1][l * 5 + ix] * hwz[l];
      }
    }
#pragma unroll
    for (int c = 0; c < 2; c++) st[le * RE2_ELEM + 2 * p + c] = -(wz * t1[c] + wx * t2[c]);
  }
  __syncthreads();
  const int nel = nspec - first < NE ? nspec - first : NE;
  float4 *dst = reinterpret_cast<float4 *>(re2 + static_cast<size_t>(first) * RE2_ELEM);
  const float4 *from = reinterpret_cast<const float4 *>(st);
  for (int qv = t; qv < nel * (RE2_ELEM / 4); qv += 128) dst[qv] = from[qv];
}

// One thread per element point on the absorbing boundary (entries sorted by
// element point); its edge entries (two at a corner) are summed in order and
// added to the point's element force once.
__global__ __launch_bounds__(256) void stacey_re2_kernel(int nseg, const int *__restrict__ seg,
                                                         const int *__restrict__ gather, const int *__restrict__ slot,
                                                         const float *__restrict__ normal,
                                                         const float *__restrict__ weight,
                                                         const float *__restrict__ rho_vp,
                                                         const float *__restrict__ rho_vs, Fields f,
                                                         float *__restrict__ re2) {
#pragma clang fp contract(off)
  const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nseg) return;
  const int q0 = seg[sidx], q1 = seg[sidx + 1];
  const int g = gather[q0];
  const int sl = slot[q0];
  const int e = sl / 25, p = sl - (sl / 25) * 25;
  const int np = f.nglob_pad;
  const float v[2] = {f.v[g], f.v[np + g]};
  float sum[2] = {0.0f, 0.0f};
  for (int qq = q0; qq < q1; qq++) {
    const float n[2] = {normal[2 * qq], normal[2 * qq + 1]};
    const float vn = v[0] * n[0] + v[1] * n[1];
    const float nn = n[0] * n[0] + n[1] * n[1];
    const float j1 = sqrtf(nn);
    const float w = weight[qq], rvp = rho_vp[qq], rvs = rho_vs[qq];
#pragma unroll
    for (int d = 0; d < 2; d++) {
      const float fac = ((vn * n[d] / (j1 * j1)) * (rvp - rvs)) + v[d] * rvs;
      sum[d] += -1.0f * fac * j1 * w;
    }
  }
  float *o = re2 + static_cast<size_t>(e) * RE2_ELEM + 2 * p;
  o[0] += sum[0];
  o[1] += sum[1];
}

// Sources in application order, one thread per element point.
__global__ __launch_bounds__(64) void sources_re2_kernel(int nsrc, int istep_arg, const int *__restrict__ order,
                                                         const int *__restrict__ ispec,
                                                         const float *__restrict__ array,
                                                         const float *__restrict__ stf, float *__restrict__ re2,
                                                         const int *__restrict__ step) {
#pragma clang fp contract(off)
  const int p = threadIdx.x;
  if (p >= 25) return;
  const int istep = step ? *step : istep_arg;
  for (int k = 0; k < nsrc; k++) {
    const int s = order[k];
    const float sv = stf[static_cast<size_t>(istep) * nsrc + s];
    float *o = re2 + static_cast<size_t>(ispec[s]) * RE2_ELEM + 2 * p;
#pragma unroll
    for (int c = 0; c < 2; c++) o[c] += sv * array[(static_cast<size_t>(s) * 2 + c) * 25 + p];
  }
}

// One thread per field point: the assembled force (its element contributions
// in a fixed order), a = M^-1 f; divide_only stores a, otherwise corrector and
// predictor follow: v += dt/2 a; u += dt v + dt^2/2 a; v += dt/2 a, and a is
// not stored (nothing in the fused loop reads it; the engine recomputes it
// from re2 with divide_only when a caller asks for it).
// The first AU2_BATCH contributions of a point are loaded together (entry
// indices clamped to the point's last one, so every address is valid) and
// summed in ascending order, then any further ones one by one: the same sum
// as a plain loop, with two dependent round trips instead of one per entry.
// M^-1, u and v do not depend on the sum and are loaded first.
constexpr int AU2_BATCH = 4;

__global__ __launch_bounds__(256) void assemble_update_2d_kernel(int nglob, ForceSum2D fs, Fields f,
                                                                 NewmarkConst nc, int divide_only) {
#pragma clang fp contract(off)
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nglob) return;
  const int np = f.nglob_pad;
  float minv[2], u[2] = {0.0f, 0.0f}, v[2] = {0.0f, 0.0f};
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const size_t o = static_cast<size_t>(c) * np + g;
    minv[c] = f.minv[o];
    if (!divide_only) {
      u[c] = f.u[o];
      v[c] = f.v[o];
    }
  }
  const int k0 = fs.off[g], k1 = fs.off[g + 1];
  float sum[2] = {0.0f, 0.0f};
  if (k1 > k0) {
    int en[AU2_BATCH];
#pragma unroll
    for (int j = 0; j < AU2_BATCH; j++) en[j] = fs.ent[k0 + j < k1 ? k0 + j : k1 - 1];
    float2 r[AU2_BATCH];
#pragma unroll
    for (int j = 0; j < AU2_BATCH; j++) r[j] = *reinterpret_cast<const float2 *>(fs.re2 + en[j]);
#pragma unroll
    for (int j = 0; j < AU2_BATCH; j++)
      if (k0 + j < k1) {
        sum[0] += r[j].x;
        sum[1] += r[j].y;
      }
    for (int k = k0 + AU2_BATCH; k < k1; k++) {
      const float2 rr = *reinterpret_cast<const float2 *>(fs.re2 + fs.ent[k]);
      sum[0] += rr.x;
      sum[1] += rr.y;
    }
  }
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const size_t o = static_cast<size_t>(c) * np + g;
    const float a = sum[c] * minv[c];
    if (divide_only) {
      f.a[o] = a;
      continue;
    }
    const float v1 = v[c] + nc.dto2 * a;
    f.u[o] = u[c] + (nc.dt * v1 + nc.dt2o2 * a);
    f.v[o] = v1 + nc.dto2 * a;
  }
}

// Adjoint and backward fields together (the combined loop's two
// update_wavefields calls, time_marching.tpp:160-186): one pass over the
// elements loads the geometry and properties once and computes both fields'
// forces.  The backward field replays the recorded force of Stacey elements
// (compute_stiffness_interaction.tpp:119-136) and adds nothing when
// bwd_active is 0.  With frechet set, the strain parts of the kernels
// (frechet_derivative.hpp:58-118: mu and kappa, each times its coefficient and
// dt) are accumulated at every element point from the same gradients: both
// displacements are the fields' values after this step.
template <bool ATOMIC>
__global__ __launch_bounds__(128) void stiffness_pair_2d_kernel(
    int first, int count, const int *__restrict__ gather, const float *__restrict__ xix_, const float *__restrict__ xiz_,
    const float *__restrict__ gammax_, const float *__restrict__ gammaz_, const float *__restrict__ jac_,
    const float *__restrict__ kappa_, const float *__restrict__ mu_, GllConst q, Fields adj, Fields bwd, int bwd_active,
    const int *__restrict__ rec_index, const float *__restrict__ rec, int frechet, float dt,
    float *__restrict__ Kmu, float *__restrict__ Kkappa, const int *__restrict__ step, long long rec_stride) {
  __shared__ float su[2][E2][2][25];  // [field][element][comp][point]
  __shared__ float sF[2][E2][4][25];

  const int t = threadIdx.x;
  const int el = t / 25;
  const int p = t - el * 25;
  const int iz = p / 5, ix = p - (p / 5) * 5;
  const int el_idx = blockIdx.x * E2 + el;
  const bool active = (el < E2) && (el_idx < count);
  // every lane loads (inactive ones re-read the range's last element; all
  // addresses valid), all before the barrier and without branches, so the
  // loads are in flight together
  const int e = first + (el_idx < count ? el_idx : count - 1);
  const size_t idx = static_cast<size_t>(e) * 25 + p;
  const int np = adj.nglob_pad;

  const int g = gather[idx];
  float ua[2], ub[2];
#pragma unroll
  for (int c = 0; c < 2; c++) {
    ua[c] = adj.u[c * np + g];
    ub[c] = bwd.u[c * np + g];
  }
  const float xix = xix_[idx], xiz = xiz_[idx], gammax = gammax_[idx], gammaz = gammaz_[idx], J = jac_[idx];
  const float kappa = kappa_[idx], mu = mu_[idx];
  float hpx[5], hpz[5], hwx[5], hwz[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    hpx[l] = q.hp[ix][l];
    hpz[l] = q.hp[iz][l];
    hwx[l] = q.hpw[ix][l];
    hwz[l] = q.hpw[iz][l];
  }
  const float wx = q.w[ix], wz = q.w[iz];
  if (active) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      su[0][el][c][p] = ua[c];
      su[1][el][c][p] = ub[c];
    }
  }
  __syncthreads();

  if (active) {
    const float lp2m = kappa + (4.0f / 3.0f) * mu;
    const float lambda = lp2m - 2.0f * mu;
    float du[2][2][2];  // [field][c][j]
#pragma unroll
    for (int fi = 0; fi < 2; fi++) {
      float dxi[2] = {0.0f, 0.0f}, dga[2] = {0.0f, 0.0f};
#pragma unroll
      for (int l = 0; l < 5; l++) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
          dxi[c] += hpx[l] * su[fi][el][c][iz * 5 + l];
          dga[c] += hpz[l] * su[fi][el][c][l * 5 + ix];
        }
      }
#pragma unroll
      for (int c = 0; c < 2; c++) {
        du[fi][c][0] = xix * dxi[c] + gammax * dga[c];
        du[fi][c][1] = xiz * dxi[c] + gammaz * dga[c];
      }
      const float sxx = lp2m * du[fi][0][0] + lambda * du[fi][1][1];
      const float szz = lp2m * du[fi][1][1] + lambda * du[fi][0][0];
      const float sxz = mu * (du[fi][0][1] + du[fi][1][0]);
      sF[fi][el][0][p] = J * (sxx * xix + sxz * xiz);
      sF[fi][el][1][p] = J * (sxx * gammax + sxz * gammaz);
      sF[fi][el][2][p] = J * (sxz * xix + szz * xiz);
      sF[fi][el][3][p] = J * (sxz * gammax + szz * gammaz);
    }
    if (frechet) {
      const float kap2 = lp2m - mu;
      const float ad_dsxx = du[0][0][0];
      const float ad_dsxz = 0.5f * (du[0][1][0] + du[0][0][1]);
      const float ad_dszz = du[0][1][1];
      const float b_dsxx = du[1][0][0];
      const float b_dsxz = 0.5f * (du[1][1][0] + du[1][0][1]);
      const float b_dszz = du[1][1][1];
      const float kappa_kl = (ad_dsxx + ad_dszz) * (b_dsxx + b_dszz);
      const float mu_kl =
          (ad_dsxx * b_dsxx + ad_dszz * b_dszz + 2.0f * ad_dsxz * b_dsxz - (1.0f / 3.0f) * kappa_kl);
      const size_t kidx = static_cast<size_t>(e) * 25 + p;
      Kkappa[kidx] += -1.0f * kap2 * dt * kappa_kl;
      Kmu[kidx] += -2.0f * mu * dt * mu_kl;
    }
  }
  __syncthreads();
  if (!active) return;
#pragma unroll
  for (int fi = 0; fi < 2; fi++) {
    if (fi == 1 && !bwd_active) break;
    float t1[2] = {0.0f, 0.0f}, t2[2] = {0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < 5; l++) {
#pragma unroll
      for (int c = 0; c < 2; c++) {
        t1[c] += sF[fi][el][2 * c][iz * 5 + l] * hwx[l];
        t2[c] += sF[fi][el][2 * c + 1][l * 5 + ix] * hwz[l];
      }
    }
    float r[2];
#pragma unroll
    for (int c = 0; c < 2; c++) r[c] = -(wz * t1[c] + wx * t2[c]);
    if (fi == 1) {
      const int ri = rec_index[e];
      if (ri >= 0) {
        const float *rp = rec + (step ? *step * rec_stride : 0) + (static_cast<size_t>(ri) * 25 + p) * 2;
        r[0] = rp[0];
        r[1] = rp[1];
      }
    }
    Fields &f = fi == 0 ? adj : bwd;
#pragma unroll
    for (int c = 0; c < 2; c++) scatter_add<ATOMIC>(&f.a[c * np + g], r[c]);
  }
}

// End of a combined step for both fields (float4 over points): mass division
// and corrector (newmark.tpp:11-56; the backward with -dt/2, :204-222), the
// rho kernel's pointwise product R += dt a+ . u_b (frechet_derivative.hpp:96-110;
// a+ after the division, u_b this step's backward displacement), then the
// predictors of the next step (+dt for the adjoint, -dt for the backward,
// newmark.tpp:225-242).  bwd_correct = 0 when the backward state is already
// corrected (the first step, where it is the forward run's final state).
__global__ __launch_bounds__(256) void update_pair_kernel(int n4, int ncomp, Fields adj, Fields bwd, NewmarkConst nc,
                                                          int bwd_correct, float *__restrict__ R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  const size_t n4s = adj.nglob_pad / 4;
  float4 rho_acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  for (int c = 0; c < ncomp; c++) {
    const size_t o = static_cast<size_t>(c) * n4s + i;
    const float4 m = reinterpret_cast<const float4 *>(adj.minv)[o];
    float4 a = reinterpret_cast<const float4 *>(adj.a)[o];
    float4 u = reinterpret_cast<const float4 *>(adj.u)[o];
    float4 v = reinterpret_cast<const float4 *>(adj.v)[o];
    float4 ab = reinterpret_cast<const float4 *>(bwd.a)[o];
    float4 ub = reinterpret_cast<const float4 *>(bwd.u)[o];
    float4 vb = reinterpret_cast<const float4 *>(bwd.v)[o];
#define SFX_PAIR(X)                                          \
  {                                                          \
    a.X = a.X * m.X;                                         \
    v.X = v.X + nc.dto2 * a.X;                               \
    if (bwd_correct) {                                       \
      ab.X = ab.X * m.X;                                     \
      vb.X = vb.X + (-nc.dto2) * ab.X;                       \
    }                                                        \
    rho_acc.X += a.X * ub.X;                                 \
    u.X = u.X + (nc.dt * v.X + nc.dt2o2 * a.X);              \
    v.X = v.X + nc.dto2 * a.X;                               \
    ub.X = ub.X + (-nc.dt * vb.X + nc.dt2o2 * ab.X);         \
    vb.X = vb.X + (-nc.dto2) * ab.X;                         \
  }
    SFX_PAIR(x) SFX_PAIR(y) SFX_PAIR(z) SFX_PAIR(w)
#undef SFX_PAIR
    reinterpret_cast<float4 *>(adj.u)[o] = u;
    reinterpret_cast<float4 *>(adj.v)[o] = v;
    reinterpret_cast<float4 *>(adj.a)[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    reinterpret_cast<float4 *>(bwd.u)[o] = ub;
    reinterpret_cast<float4 *>(bwd.v)[o] = vb;
    reinterpret_cast<float4 *>(bwd.a)[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  float4 rr = reinterpret_cast<float4 *>(R)[i];
  rr.x += nc.dt * rho_acc.x;
  rr.y += nc.dt * rho_acc.y;
  rr.z += nc.dt * rho_acc.z;
  rr.w += nc.dt * rho_acc.w;
  reinterpret_cast<float4 *>(R)[i] = rr;
}

// R += dt a+ . u_b on corrected states (checkpoint steps).
__global__ __launch_bounds__(256) void rho_product_kernel(int n, int ncomp, Fields adj, Fields bwd, float dt,
                                                          float *__restrict__ R) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int c = 0; c < ncomp; c++) s += adj.a[static_cast<size_t>(c) * adj.nglob_pad + i] *
                                      bwd.u[static_cast<size_t>(c) * bwd.nglob_pad + i];
  R[i] += dt * s;
}

// The six kernels from the accumulated mu and kappa kernels and R
// (frechet_derivative.hpp:110-135): rho = -rho_p R(g), rhop = rho + kappa + mu,
// beta = 2 (mu - 4/3 mu_p / kappa_p kappa), alpha = 2 (1 + 4/3 mu_p / kappa_p) kappa,
// with kappa_p = lambda + mu.  out (6, nspec, 25).
__global__ __launch_bounds__(256) void frechet_finish_kernel(int npts, const int *__restrict__ gather,
                                                             const float *__restrict__ kappa_,
                                                             const float *__restrict__ mu_,
                                                             const float *__restrict__ rho_,
                                                             const float *__restrict__ R,
                                                             const float *__restrict__ Kmu,
                                                             const float *__restrict__ Kkappa,
                                                             float *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const float mu = mu_[i], rho = rho_[i];
  const float kap2 = kappa_[i] + (4.0f / 3.0f) * mu - mu;
  const float rho_kl = -1.0f * rho * R[gather[i]];
  const float mu_kl = Kmu[i], kappa_kl = Kkappa[i];
  out[i] = rho_kl;
  out[npts + i] = mu_kl;
  out[2 * static_cast<size_t>(npts) + i] = kappa_kl;
  out[3 * static_cast<size_t>(npts) + i] = rho_kl + kappa_kl + mu_kl;
  out[4 * static_cast<size_t>(npts) + i] = 2.0f * (1.0f + (4.0f / 3.0f) * mu / kap2) * kappa_kl;
  out[5 * static_cast<size_t>(npts) + i] = 2.0f * (mu_kl - (4.0f / 3.0f) * mu / kap2 * kappa_kl);
}

// Fréchet derivatives of one step (compute_material_derivatives.tpp:17-154,
// include/medium/dim2/elastic/isotropic/frechet_derivative.hpp:14-138):
// gradients of the adjoint and backward displacements per element point,
// kernels K[k][slot * 25 + p] += (rho, mu, kappa, rhop, alpha, beta).
__global__ __launch_bounds__(128) void frechet_2d_kernel(
    int nspec, const int *__restrict__ gather, const float *__restrict__ xix_, const float *__restrict__ xiz_,
    const float *__restrict__ gammax_, const float *__restrict__ gammaz_, const float *__restrict__ kappa_,
    const float *__restrict__ mu_, const float *__restrict__ rho_, GllConst q, Fields adj, Fields bwd, float dt,
    float *__restrict__ K) {
  __shared__ float sa[E2][2][25];
  __shared__ float sb[E2][2][25];
  const int t = threadIdx.x;
  const int el = t / 25;
  const int p = t - el * 25;
  const int iz = p / 5, ix = p - (p / 5) * 5;
  const int e = blockIdx.x * E2 + el;
  const bool active = (el < E2) && (e < nspec);
  const size_t idx = static_cast<size_t>(e) * 25 + p;
  const int np = adj.nglob_pad;
  int g = 0;
  if (active) {
    g = gather[idx];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      sa[el][c][p] = adj.u[c * np + g];
      sb[el][c][p] = bwd.u[c * np + g];
    }
  }
  __syncthreads();
  if (!active) return;
  const float xix = xix_[idx], xiz = xiz_[idx], gammax = gammax_[idx], gammaz = gammaz_[idx];
  float df[2][2], dg[2][2];  // d u_c / d x_j (gradient.hpp:81-103)
#pragma unroll
  for (int c = 0; c < 2; c++) {
    float fxi = 0.0f, fga = 0.0f, gxi = 0.0f, gga = 0.0f;
#pragma unroll
    for (int l = 0; l < 5; l++) {
      fxi += q.hp[ix][l] * sa[el][c][iz * 5 + l];
      fga += q.hp[iz][l] * sa[el][c][l * 5 + ix];
      gxi += q.hp[ix][l] * sb[el][c][iz * 5 + l];
      gga += q.hp[iz][l] * sb[el][c][l * 5 + ix];
    }
    df[c][0] = xix * fxi + gammax * fga;
    df[c][1] = xiz * fxi + gammaz * fga;
    dg[c][0] = xix * gxi + gammax * gga;
    dg[c][1] = xiz * gxi + gammaz * gga;
  }
  const float mu = mu_[idx], rho = rho_[idx];
  const float lp2m = kappa_[idx] + (4.0f / 3.0f) * mu;
  const float kappa = lp2m - mu;
  const float ad_dsxx = df[0][0];
  const float ad_dsxz = 0.5f * (df[1][0] + df[0][1]);
  const float ad_dszz = df[1][1];
  const float b_dsxx = dg[0][0];
  const float b_dsxz = 0.5f * (dg[1][0] + dg[0][1]);
  const float b_dszz = dg[1][1];
  float kappa_kl = (ad_dsxx + ad_dszz) * (b_dsxx + b_dszz);
  float mu_kl = (ad_dsxx * b_dsxx + ad_dszz * b_dszz + 2.0f * ad_dsxz * b_dsxz - (1.0f / 3.0f) * kappa_kl);
  float rho_kl = adj.a[g] * bwd.u[g] + adj.a[np + g] * bwd.u[np + g];
  kappa_kl = -1.0f * kappa * dt * kappa_kl;
  mu_kl = -2.0f * mu * dt * mu_kl;
  rho_kl = -1.0f * rho * dt * rho_kl;
  const float rhop_kl = rho_kl + kappa_kl + mu_kl;
  const float beta_kl = 2.0f * (mu_kl - (4.0f / 3.0f) * mu / kappa * kappa_kl);
  const float alpha_kl = 2.0f * (1.0f + (4.0f / 3.0f) * mu / kappa) * kappa_kl;
  const size_t n = static_cast<size_t>(nspec) * 25;
  K[idx] += rho_kl;
  K[n + idx] += mu_kl;
  K[2 * n + idx] += kappa_kl;
  K[3 * n + idx] += rhop_kl;
  K[4 * n + idx] += alpha_kl;
  K[5 * n + idx] += beta_kl;
}

// ---------------------------------------------------------------------------
// 3-D stiffness on regular hexahedra.  One thread owns a (i, j) column of an
// element and keeps its 5 points along k in registers: the gamma-direction
// contractions never touch LDS; xi and eta contractions read the element's
// displacement and stress integrand from LDS.  5 elements per 128-thread block
// (125 active threads).
// ---------------------------------------------------------------------------
constexpr int E3 = 5;

template <bool ATOMIC>
__global__ __launch_bounds__(128) void stiffness_3d_kernel(int first, int count, const int *__restrict__ gather,
                                                           const float *__restrict__ kappa_,
                                                           const float *__restrict__ mu_, float xix, float ety,
                                                           float gaz, float jac, GllConst q, Fields f) {
  __shared__ float su[E3][3][125];
  __shared__ float sF0[E3][3][125];  // J xix sigma(c, x)
  __shared__ float sF1[E3][3][125];  // J ety sigma(c, y)

  const int t = threadIdx.x;
  const int el = t / 25;
  const int col = t - el * 25;
  const int j = col / 5, i = col - (col / 5) * 5;
  const int el_idx = blockIdx.x * E3 + el;
  const bool active = (el < E3) && (el_idx < count);
  const int e = first + el_idx;
  const size_t base = static_cast<size_t>(e) * 125 + col;
  const int np = f.nglob_pad;

  float hpi[5], hpj[5], hwi[5], hwj[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    hpi[l] = q.hp[i][l];
    hpj[l] = q.hp[j][l];
    hwi[l] = q.hpw[i][l];
    hwj[l] = q.hpw[j][l];
  }
  const float wi = q.w[i], wj = q.w[j];

  int g[5];
  float u[5][3];
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++) g[k] = gather[base + k * 25];
#pragma unroll
    for (int k = 0; k < 5; k++) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        u[k][c] = f.u[c * np + g[k]];
        su[el][c][k * 25 + col] = u[k][c];
      }
    }
  }
  __syncthreads();

  float F2[5][3];  // J gaz sigma(c, z), kept in registers
  if (active) {
    const float sx = jac * xix, sy = jac * ety, sz = jac * gaz;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      float dxi[3] = {0, 0, 0}, det[3] = {0, 0, 0}, dga[3] = {0, 0, 0};
#pragma unroll
      for (int l = 0; l < 5; l++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          dxi[c] += hpi[l] * su[el][c][k * 25 + j * 5 + l];
          det[c] += hpj[l] * su[el][c][k * 25 + l * 5 + i];
          dga[c] += q.hp[k][l] * u[l][c];
        }
      }
      const float kappa = kappa_[base + k * 25];
      const float mu = mu_[base + k * 25];
      const float lp2m = kappa + (4.0f / 3.0f) * mu;
      const float lambda = lp2m - 2.0f * mu;
      const float exx = xix * dxi[0], eyy = ety * det[1], ezz = gaz * dga[2];
      const float sxx = lp2m * exx + lambda * (eyy + ezz);
      const float syy = lp2m * eyy + lambda * (exx + ezz);
      const float szz = lp2m * ezz + lambda * (exx + eyy);
      const float sxy = mu * (ety * det[0] + xix * dxi[1]);
      const float sxz = mu * (gaz * dga[0] + xix * dxi[2]);
      const float syz = mu * (gaz * dga[1] + ety * det[2]);
      const int p = k * 25 + col;
      sF0[el][0][p] = sx * sxx;
      sF0[el][1][p] = sx * sxy;
      sF0[el][2][p] = sx * sxz;
      sF1[el][0][p] = sy * sxy;
      sF1[el][1][p] = sy * syy;
      sF1[el][2][p] = sy * syz;
      F2[k][0] = sz * sxz;
      F2[k][1] = sz * syz;
      F2[k][2] = sz * szz;
    }
  }
  __syncthreads();
  if (!active) return;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    float r0[3] = {0, 0, 0}, r1[3] = {0, 0, 0}, r2[3] = {0, 0, 0};
#pragma unroll
    for (int l = 0; l < 5; l++) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        r0[c] += sF0[el][c][k * 25 + j * 5 + l] * hwi[l];
        r1[c] += sF1[el][c][k * 25 + l * 5 + i] * hwj[l];
        r2[c] += F2[l][c] * q.hpw[k][l];
      }
    }
    const float wjk = wj * q.w[k], wik = wi * q.w[k], wij = wi * wj;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float r = -(wjk * r0[c] + wik * r1[c] + wij * r2[c]);
      scatter_add<ATOMIC>(&f.a[c * np + g[k]], r);
    }
  }
}

// ---------------------------------------------------------------------------
// Stacey absorbing traction on a flat list of boundary element points
// (stacey.tpp:36-72): t = -[(v.n / |n|^2) n (rho vp - rho vs) + v rho vs] |n| w
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void stacey_kernel(int nbnd, const int *__restrict__ gather,
                                                     const float *__restrict__ normal,
                                                     const float *__restrict__ weight,
                                                     const float *__restrict__ rho_vp,
                                                     const float *__restrict__ rho_vs, Fields f,
                                                     const int *__restrict__ bnd_rec, float *__restrict__ rec,
                                                     const int *__restrict__ step, long long rec_stride) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nbnd) return;
  const int g = gather[q];
  const int np = f.nglob_pad;
  float n[D], v[D];
  float vn = 0.0f, nn = 0.0f;
#pragma unroll
  for (int d = 0; d < D; d++) {
    n[d] = normal[q * D + d];
    v[d] = f.v[d * np + g];
    vn += v[d] * n[d];
    nn += n[d] * n[d];
  }
  const float j1 = sqrtf(nn);
  const float w = weight[q], rvp = rho_vp[q], rvs = rho_vs[q];
#pragma unroll
  for (int d = 0; d < D; d++) {
    const float fac = ((vn * n[d] / (j1 * j1)) * (rvp - rvs)) + v[d] * rvs;
    const float tr = -1.0f * fac * j1 * w;
    atomic_add(&f.a[d * np + g], tr);
    if (rec) rec[(step ? *step * rec_stride : 0) + static_cast<size_t>(bnd_rec[q]) * D + d] += tr;
  }
}

// ---------------------------------------------------------------------------
// Sources: a += stf(istep, s) * source_array(s, c, p); one workgroup per
// source (in application order), one thread per element point.  Sources
// sharing a global point add with float atomics, as the reference's
// atomic_add_on_device does.
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(128) void sources_kernel(int npe, int nsrc, int istep_arg, const int *__restrict__ order,
                                                      const int *__restrict__ ispec, const int *__restrict__ gather,
                                                      const float *__restrict__ array,
                                                      const float *__restrict__ stf, Fields f,
                                                      const int *__restrict__ step, int norder) {
  const int p = threadIdx.x;
  const int k = blockIdx.x;
  if (p >= npe || k >= norder) return;
  const int istep = step ? *step : istep_arg;
  const int s = order[k];
  const float sv = stf[static_cast<size_t>(istep) * nsrc + s];
  const int g = gather[static_cast<size_t>(ispec[s]) * npe + p];
#pragma unroll
  for (int c = 0; c < D; c++)
    atomic_add(&f.a[c * f.nglob_pad + g], sv * array[(static_cast<size_t>(s) * D + c) * npe + p]);
}

// ---------------------------------------------------------------------------
// Seismograms: one block per receiver, Lagrange interpolation of the
// element field, block reduction.
// ---------------------------------------------------------------------------
// Pressure at element point p of 2-D element e (include/medium/dim2/elastic/
// isotropic/wavefield.hpp:49-85): -(lambda + 2/3 mu) div u from the element's
// displacement (su, staged in LDS), the per-point derivatives and properties;
// the factor is formed in double as the reference's double constant makes it.
__device__ __forceinline__ float pressure_2d(const PointGeom2D &pg, const GllConst &q, const float (*su)[25], int e,
                                             int p) {
  const int iz = p / 5, ix = p - (p / 5) * 5;
  float dxi[2] = {0.0f, 0.0f}, dga[2] = {0.0f, 0.0f};
#pragma unroll
  for (int l = 0; l < 5; l++) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      dxi[c] += q.hp[ix][l] * su[c][iz * 5 + l];
      dga[c] += q.hp[iz][l] * su[c][l * 5 + ix];
    }
  }
  const size_t idx = static_cast<size_t>(e) * 25 + p;
  const float du00 = pg.xix[idx] * dxi[0] + pg.gammax[idx] * dga[0];
  const float du11 = pg.xiz[idx] * dxi[1] + pg.gammaz[idx] * dga[1];
  const float mu = pg.mu[idx];
  const float lp2m = pg.kappa[idx] + (4.0f / 3.0f) * mu;
  const float lambda = lp2m - 2.0f * mu;
  const double fac = static_cast<double>(lambda) + (2.0 / 3.0) * static_cast<double>(mu);
  return static_cast<float>(-1.0 * (fac * static_cast<double>(du00 + du11)));
}

template <int D>
__global__ __launch_bounds__(128) void seismogram_kernel(int npe, int nrec, int ntypes, const int *__restrict__ types,
                                                         int isig, const int *__restrict__ ispec,
                                                         const int *__restrict__ gather, const float *__restrict__ h,
                                                         Fields f, float dto2, int corrected,
                                                         float *__restrict__ seismo, PointGeom2D pg, GllConst q,
                                                         const int *__restrict__ step, int nsb, int sig_top,
                                                         ForceSum2D fs, const int *__restrict__ rec_list) {
  __shared__ float red[2][D][64];
  __shared__ float su[2][25];
  if (step) {  // sample only on steps istep % nsb == 0 (the same for the whole block)
    const int is = *step;
    if (is % nsb != 0) return;
    isig = sig_top >= 0 ? sig_top - is / nsb : is / nsb;  // sig_top: samples counted down in time
  }
  const int r = rec_list ? rec_list[blockIdx.x] : blockIdx.x;
  const int p = threadIdx.x;
  const int np = f.nglob_pad;
  int g = 0;
  float hw = 0.0f;
  if (p < npe) {
    g = gather[static_cast<size_t>(ispec[r]) * npe + p];
    hw = h[static_cast<size_t>(r) * npe + p];
  }
  for (int it = 0; it < ntypes; it++) {
    const int type = types[it];
    float val[D];
    if (D == 2 && type == 3) {  // pressure: one value, second component 0
      if (p < 25) {
        su[0][p] = f.u[g];
        su[1][p] = f.u[np + g];
      }
      __syncthreads();
      val[0] = p < 25 ? hw * pressure_2d(pg, q, su, ispec[r], p) : 0.0f;
#pragma unroll
      for (int c = 1; c < D; c++) val[c] = 0.0f;
    } else
#pragma unroll
    for (int c = 0; c < D; c++) {
      float x = 0.0f;
      if (p < npe) {
        if (type == 0) {
          x = f.u[c * np + g];
        } else {
          float acc;
          if (D == 2 && fs.re2) {  // gather path: the assembled force, summed as the update sums it
            acc = 0.0f;
            for (int k = fs.off[g]; k < fs.off[g + 1]; k++) acc += fs.re2[fs.ent[k] + c];
          } else {
            acc = f.a[c * np + g];
          }
          if (corrected) acc = acc * f.minv[c * np + g];
          if (type == 1) {
            x = f.v[c * np + g];
            // the gather path rounds v + dt/2 a as its update does (no fused
            // multiply-add)
            if (corrected) x = fs.re2 ? plus_product_rounded(x, dto2, acc) : x + dto2 * acc;
          } else {
            x = acc;
          }
        }
      }
      val[c] = hw * x;
    }
    // wave reduction then across the two waves
#pragma unroll
    for (int c = 0; c < D; c++) {
      float s = val[c];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
      if ((p & 63) == 0) red[0][c][p >> 6] = s;
    }
    __syncthreads();
    if (p == 0) {
#pragma unroll
      for (int c = 0; c < D; c++)
        seismo[((static_cast<size_t>(isig) * ntypes + it) * nrec + r) * D + c] = red[0][c][0] + red[0][c][1];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fused mass division + corrector(n) + predictor(n+1), 4 points per thread.
//   a <- a M^-1; v <- v + dt/2 a            (divide_mass_matrix, corrector)
//   u <- u + (dt v + dt^2/2 a); v <- v + dt/2 a; a <- 0   (predictor)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void update_fused_kernel(int n4, int ncomp, Fields f, NewmarkConst nc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  for (int c = 0; c < ncomp; c++) {
    const size_t o = static_cast<size_t>(c) * (f.nglob_pad / 4) + i;
    float4 a = reinterpret_cast<const float4 *>(f.a)[o];
    const float4 m = reinterpret_cast<const float4 *>(f.minv)[o];
    float4 u = reinterpret_cast<const float4 *>(f.u)[o];
    float4 v = reinterpret_cast<const float4 *>(f.v)[o];
#define SFX_UPD(X)                                   \
  {                                                  \
    const float ac = a.X * m.X;                      \
    const float v1 = v.X + nc.dto2 * ac;             \
    u.X = u.X + (nc.dt * v1 + nc.dt2o2 * ac);        \
    v.X = v1 + nc.dto2 * ac;                         \
  }
    SFX_UPD(x) SFX_UPD(y) SFX_UPD(z) SFX_UPD(w)
#undef SFX_UPD
    reinterpret_cast<float4 *>(f.u)[o] = u;
    reinterpret_cast<float4 *>(f.v)[o] = v;
    reinterpret_cast<float4 *>(f.a)[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// Predictor alone (newmark.tpp:60-116).
// ROUNDED: every product and sum rounded on its own (no fused multiply-add),
// as the 2-D gather path's fused update forms them.
template <bool ROUNDED>
__global__ __launch_bounds__(256) void predictor_kernel(int n, int ncomp, Fields f, NewmarkConst nc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < ncomp; c++) {
    const size_t o = static_cast<size_t>(c) * f.nglob_pad + i;
    const float a = f.a[o], v = f.v[o];
    if constexpr (ROUNDED) {
#pragma clang fp contract(off)
      f.u[o] = f.u[o] + (nc.dt * v + nc.dt2o2 * a);
      f.v[o] = v + nc.dto2 * a;
    } else {
      f.u[o] = f.u[o] + (nc.dt * v + nc.dt2o2 * a);
      f.v[o] = v + nc.dto2 * a;
    }
    f.a[o] = 0.0f;
  }
}

// Mass division (divide_mass_matrix.tpp:42-49) and/or corrector (newmark.tpp:11-56).
template <bool ROUNDED>
__global__ __launch_bounds__(256) void divide_corrector_kernel(int n, int ncomp, Fields f, NewmarkConst nc,
                                                               int divide, int correct) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < ncomp; c++) {
    const size_t o = static_cast<size_t>(c) * f.nglob_pad + i;
    float a = f.a[o];
    if (divide) {
      a = a * f.minv[o];
      f.a[o] = a;
    }
    if (correct) {
      if constexpr (ROUNDED) {
#pragma clang fp contract(off)
        f.v[o] = f.v[o] + nc.dto2 * a;
      } else {
        f.v[o] = f.v[o] + nc.dto2 * a;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 3-D structured path with buffered element forces.
//
// elemforce_3d: the stiffness of stiffness_3d_kernel, but the element's
// -K_e u_e is stored to re (element-major, coalesced) instead of scattered.
// Displacement addresses come from the tiled lattice position; the
// workgroup -> element-block mapping is XCD-aware: the 8 XCDs receive
// workgroups round-robin, so each XCD is given one contiguous slab of
// elements and its L2 sees the shared faces of neighbouring elements.
// ---------------------------------------------------------------------------
// EPROP: kappa and mu are read once per element from kmu (kappa, mu pairs)
// instead of at each of its 125 points; the values are the same bits.
template <bool EPROP>
__global__ __launch_bounds__(128) void elemforce_3d_kernel(Lattice lat, const float *__restrict__ kappa_,
                                                           const float *__restrict__ mu_,
                                                           const float *__restrict__ kmu, Gll3 q,
                                                           const float *__restrict__ uf, int np,
                                                           double inv_nx, double inv_nxy,
                                                           float *__restrict__ re) {
  // LDS: the displacement, then the forces in slot order, at su[el * SU_EL +
  // c * RE_STRIDE + x]; the five distinct stresses sxx, sxy, sxz, syy, syz (sxy
  // once: it is both sigma(x, y) and sigma(y, x)) at sS[el * SS_EL + m * 125 + p].
  // The element strides are padded so that the three elements a wave touches
  // fall on different LDS banks (392 = 8 and 632 = 24 mod 32, 8 and 56 mod 64):
  // the row reads of the contractions are then free of bank conflicts.  20.4 KB
  // per workgroup lets 8 workgroups share a CU's LDS.
  constexpr int SU_EL = RE_ELEM + 8, SS_EL = 5 * 125 + 7;
  __shared__ __align__(16) float su[(E3 - 1) * SU_EL + RE_ELEM];
  __shared__ float sS[(E3 - 1) * SS_EL + 5 * 125];

  const int nspec = lat.nx * lat.ny * lat.nz;
  const int t = threadIdx.x;
  const int el = t / 25;
  const int col = t - el * 25;
  const int j = col / 5, i = col - (col / 5) * 5;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int e = blk * E3 + el;
  const bool active = (el < E3) && (e < nspec);
  const size_t base = static_cast<size_t>(e) * 125 + col;

  float hpi[5], hpj[5], hwi[5], hwj[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    hpi[l] = q.hx[i][l];
    hpj[l] = q.hy[j][l];
    hwi[l] = q.wx[i][l];
    hwj[l] = q.wy[j][l];
  }
  const float wi = q.w[i], wj = q.w[j];

  // material first, so its latency overlaps the displacement loads
  float kap[5], muv[5];
  if (active) {
    if constexpr (EPROP) {
      const float ke = kmu[2 * static_cast<size_t>(e)], me = kmu[2 * static_cast<size_t>(e) + 1];
#pragma unroll
      for (int k = 0; k < 5; k++) {
        kap[k] = ke;
        muv[k] = me;
      }
    } else {
#pragma unroll
      for (int k = 0; k < 5; k++) {
        kap[k] = kappa_[base + k * 25];
        muv[k] = mu_[base + k * 25];
      }
    }
  }
  float u[5][3];
  if (active) {
    // e / nx and e / (nx ny) as floor((e + 1/2) / d) in FP64: the product's
    // error (below 2^-51 relative) stays far from the 1/(2d) margin to the next
    // integer for any int e, and it costs a few instructions instead of an
    // integer division sequence
    const int eyz = static_cast<int>((e + 0.5) * inv_nx), ez = static_cast<int>((e + 0.5) * inv_nxy);
    const int ex = e - eyz * lat.nx, ey = eyz - ez * lat.ny;
    // one 32-bit point offset per layer against the three (uniform) component
    // bases, so the loads need no 64-bit address arithmetic per component
    const float *__restrict__ ub[3] = {uf, uf + np, uf + 2 * static_cast<size_t>(np)};
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const unsigned d = static_cast<unsigned>(tiled_index(lat, ex, ey, ez, i, j, k));
#pragma unroll
      for (int c = 0; c < 3; c++) {
        u[k][c] = ub[c][d];
        su[el * SU_EL + c * RE_STRIDE + k * 25 + col] = u[k][c];
      }
    }
  }
  __syncthreads();

  float F2[5][3];
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      // strain components from the scaled contractions: d[c] = du_c / dx etc.
      float dx[3] = {0, 0, 0}, dy[3] = {0, 0, 0}, dz[3] = {0, 0, 0};
#pragma unroll
      for (int l = 0; l < 5; l++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          dx[c] += hpi[l] * su[el * SU_EL + c * RE_STRIDE + k * 25 + j * 5 + l];
          dy[c] += hpj[l] * su[el * SU_EL + c * RE_STRIDE + k * 25 + l * 5 + i];
          dz[c] += q.hz[k][l] * u[l][c];
        }
      }
      const float kappa = kap[k];
      const float mu = muv[k];
      const float lp2m = kappa + (4.0f / 3.0f) * mu;
      const float lambda = lp2m - 2.0f * mu;
      const float exx = dx[0], eyy = dy[1], ezz = dz[2];
      const float sxx = lp2m * exx + lambda * (eyy + ezz);
      const float syy = lp2m * eyy + lambda * (exx + ezz);
      const float szz = lp2m * ezz + lambda * (exx + eyy);
      const float sxy = mu * (dy[0] + dx[1]);
      const float sxz = mu * (dz[0] + dx[2]);
      const float syz = mu * (dz[1] + dy[2]);
      const int p = k * 25 + col;
      sS[el * SS_EL + 0 * 125 + p] = sxx;
      sS[el * SS_EL + 1 * 125 + p] = sxy;
      sS[el * SS_EL + 2 * 125 + p] = sxz;
      sS[el * SS_EL + 3 * 125 + p] = syy;
      sS[el * SS_EL + 4 * 125 + p] = syz;
      F2[k][0] = sxz;
      F2[k][1] = syz;
      F2[k][2] = szz;
    }
  }
  __syncthreads();
  // element forces into LDS in slot order (the displacement there is dead)
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      float r0[3] = {0, 0, 0}, r1[3] = {0, 0, 0}, r2[3] = {0, 0, 0};
#pragma unroll
      for (int l = 0; l < 5; l++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          // row c of sigma along x (sxx, sxy, sxz) and along y (sxy, syy, syz)
          r0[c] += sS[el * SS_EL + c * 125 + k * 25 + j * 5 + l] * hwi[l];
          r1[c] += sS[el * SS_EL + (c == 0 ? 1 : c + 2) * 125 + k * 25 + l * 5 + i] * hwj[l];
          r2[c] += F2[l][c] * q.wz[k][l];
        }
      }
      // the weights negated: -(a r0 + b r1 + c r2) = (-a) r0 + (-b) r1 + (-c) r2
      // exactly (rounding is symmetric), without a sign flip per component
      const float wjk = -(wj * q.w[k]), wik = -(wi * q.w[k]), wij = -(wi * wj);
      // re_slot(i, j, k) written as selects on the thread's (i, j) and the
      // unrolled k: the switch of re_slot compiles to branches here
      const bool fi = i == 4, fj = j == 4;
      const int kb = !fi && !fj ? j * 4 + i : !fj ? 64 + j : !fi ? 80 + i : 112;
      const int kst = !fi && !fj ? 16 : fi && fj ? 1 : 4;
      const int b4 = !fi && !fj ? 96 + j * 4 + i : !fj ? 116 + j : !fi ? 120 + i : 124;
      const int slot = k < 4 ? kb + k * kst : b4;
#pragma unroll
      for (int c = 0; c < 3; c++) su[el * SU_EL + c * RE_STRIDE + slot] = wjk * r0[c] + wik * r1[c] + wij * r2[c];
    }
  }
  __syncthreads();
  // the workgroup's elements are consecutive in re: one coalesced copy in
  // 16-byte stores (scattered 4-byte stores cost about half the kernel)
  const int first = blk * E3;
  const int nel = nspec - first < E3 ? nspec - first : E3;
  float4 *dst = reinterpret_cast<float4 *>(re + static_cast<size_t>(first) * RE_ELEM);
  const float4 *from = reinterpret_cast<const float4 *>(su);
  for (int qv = t; qv < nel * (RE_ELEM / 4); qv += 128) {
    const int le = qv / (RE_ELEM / 4);
    dst[qv] = from[le * (SU_EL / 4) + qv - le * (RE_ELEM / 4)];
  }
}

// Sum, in a fixed order, of the element forces at lattice point
// (4 tx + i, 4 ty + j, 4 tz + k), i, j, k < 4: the element owning the tile,
// then the elements below in y, z, yz, x, xy, xz and xyz that share the point
// (the order of the 4-point row version below, so both give equal bits).
__device__ __forceinline__ void assemble_tiled(const Lattice &lat, const float *__restrict__ re, int tx, int ty,
                                               int tz, int i, int j, int k, float acc[3]) {
  bool first = true;
  auto add = [&](int ex, int ey, int ez, int slot) {
    if (ex < 0 || ey < 0 || ez < 0 || ex >= lat.nx || ey >= lat.ny || ez >= lat.nz) return;
    const float *r = re + ((static_cast<size_t>(ez) * lat.ny + ey) * lat.nx + ex) * RE_ELEM + slot;
#pragma unroll
    for (int c = 0; c < 3; c++) acc[c] = first ? r[c * RE_STRIDE] : acc[c] + r[c * RE_STRIDE];
    first = false;
  };
  acc[0] = acc[1] = acc[2] = 0.0f;
  add(tx, ty, tz, (k * 4 + j) * 4 + i);
  if (j == 0) add(tx, ty - 1, tz, 80 + k * 4 + i);
  if (k == 0) add(tx, ty, tz - 1, 96 + j * 4 + i);
  if (j == 0 && k == 0) add(tx, ty - 1, tz - 1, 120 + i);
  if (i == 0) add(tx - 1, ty, tz, 64 + k * 4 + j);
  if (i == 0 && j == 0) add(tx - 1, ty - 1, tz, 112 + k);
  if (i == 0 && k == 0) add(tx - 1, ty, tz - 1, 116 + j);
  if (i == 0 && j == 0 && k == 0) add(tx - 1, ty - 1, tz - 1, 124);
}

// Device slot d -> (tile, local) and whether it is a lattice point.
__device__ __forceinline__ bool tiled_point(const Lattice &lat, int d, int &tx, int &ty, int &tz, int &i, int &j,
                                            int &k) {
  const int tile = d >> 6, l = d & 63;
  i = l & 3;
  j = (l >> 2) & 3;
  k = l >> 4;
  tx = tile % lat.TX;
  ty = (tile / lat.TX) % lat.TY;
  tz = tile / (lat.TX * lat.TY);
  return (tx < lat.nx || i == 0) && (ty < lat.ny || j == 0) && (tz < lat.nz || k == 0);
}

// One thread per row of 4 points (i = 0..3) of a tile: 128-bit loads and
// stores of u, v, M^-1 and of the element-force rows.  Contributions a point
// does not receive are read from `zero` (an all-zero element block), so every
// load is unconditional and the thread issues them all before the first wait.
__global__ __launch_bounds__(256) void assemble_update_3d_kernel(Lattice lat, const float *__restrict__ re,
                                                                 const float *__restrict__ zero, Fields f,
                                                                 NewmarkConst nc,
                                                                 const float *__restrict__ minv1, int amask) {
  const int nrows = lat.TX * lat.TY * lat.TZ * 16;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= nrows) return;
  const int tile = row >> 4, j = row & 3, k = (row >> 2) & 3;
  const int tx = tile % lat.TX, ty = (tile / lat.TX) % lat.TY, tz = tile / (lat.TX * lat.TY);
  if ((ty == lat.ny && j > 0) || (tz == lat.nz && k > 0)) return;  // padding rows
  auto elem = [&](bool use, int ex, int ey, int ez) -> const float * {
    if (!use || ex < 0 || ey < 0 || ez < 0 || ex >= lat.nx || ey >= lat.ny || ez >= lat.nz) return zero;
    return re + ((static_cast<size_t>(ez) * lat.ny + ey) * lat.nx + ex) * RE_ELEM;
  };
  const bool j0 = j == 0, k0 = k == 0;
  const float *r_own = elem(true, tx, ty, tz);
  const float *r_y = elem(j0, tx, ty - 1, tz);
  const float *r_z = elem(k0, tx, ty, tz - 1);
  const float *r_yz = elem(j0 && k0, tx, ty - 1, tz - 1);
  const float *r_x = elem(true, tx - 1, ty, tz);
  const float *r_xy = elem(j0, tx - 1, ty - 1, tz);
  const float *r_xz = elem(k0, tx - 1, ty, tz - 1);
  const float *r_xyz = elem(j0 && k0, tx - 1, ty - 1, tz - 1);
  const int n4 = f.nglob_pad / 4;
  float4 m[3], u[3], v[3], acc[3];
  // M^-1 is the same for the three components except at absorbing-face
  // points (the Stacey mass term): tiles holding such points read all three,
  // the others one shared value
  const bool absorbing_tile = ((amask & 1) && tx == 0) || ((amask & 2) && tx == lat.nx) ||
                              ((amask & 4) && ty == 0) || ((amask & 8) && ty == lat.ny) ||
                              ((amask & 16) && tz == 0) || ((amask & 32) && tz == lat.nz);
  if (absorbing_tile) {
#pragma unroll
    for (int c = 0; c < 3; c++) m[c] = reinterpret_cast<const float4 *>(f.minv)[static_cast<size_t>(c) * n4 + row];
  } else {
    m[0] = reinterpret_cast<const float4 *>(minv1)[row];
    m[1] = m[0];
    m[2] = m[0];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t o = static_cast<size_t>(c) * n4 + row;
    u[c] = reinterpret_cast<const float4 *>(f.u)[o];
    v[c] = reinterpret_cast<const float4 *>(f.v)[o];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int cs = c * RE_STRIDE;
    const float4 a0 = *reinterpret_cast<const float4 *>(r_own + cs + (k * 4 + j) * 4);
    const float4 a1 = *reinterpret_cast<const float4 *>(r_y + cs + 80 + k * 4);
    const float4 a2 = *reinterpret_cast<const float4 *>(r_z + cs + 96 + j * 4);
    const float4 a3 = *reinterpret_cast<const float4 *>(r_yz + cs + 120);
    const float b0 = r_x[cs + 64 + k * 4 + j];
    const float b1 = r_xy[cs + 112 + k];
    const float b2 = r_xz[cs + 116 + j];
    const float b3 = r_xyz[cs + 124];
    // order: own, y, z, yz, then (point i = 0) x, xy, xz, xyz
    acc[c].x = ((((((a0.x + a1.x) + a2.x) + a3.x) + b0) + b1) + b2) + b3;
    acc[c].y = ((a0.y + a1.y) + a2.y) + a3.y;
    acc[c].z = ((a0.z + a1.z) + a2.z) + a3.z;
    acc[c].w = ((a0.w + a1.w) + a2.w) + a3.w;
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t o = static_cast<size_t>(c) * n4 + row;
#define SFX_UPD(X)                                   \
  {                                                  \
    const float a = acc[c].X * m[c].X;               \
    const float v1 = v[c].X + nc.dto2 * a;           \
    u[c].X = u[c].X + (nc.dt * v1 + nc.dt2o2 * a);   \
    v[c].X = v1 + nc.dto2 * a;                       \
  }
    SFX_UPD(x) SFX_UPD(y) SFX_UPD(z) SFX_UPD(w)
#undef SFX_UPD
    reinterpret_cast<float4 *>(f.u)[o] = u[c];
    reinterpret_cast<float4 *>(f.v)[o] = v[c];
  }
}

__global__ __launch_bounds__(256) void assemble_3d_kernel(Lattice lat, const float *__restrict__ re, Fields f,
                                                          int divide) {
  const int n = lat.TX * lat.TY * lat.TZ * 64;
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= n) return;
  int tx, ty, tz, i, j, k;
  if (!tiled_point(lat, d, tx, ty, tz, i, j, k)) return;
  float acc[3];
  assemble_tiled(lat, re, tx, ty, tz, i, j, k, acc);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t o = static_cast<size_t>(c) * f.nglob_pad + d;
    f.a[o] = divide ? acc[c] * f.minv[o] : acc[c];
  }
}

// One thread per element point adds every source in source order: two
// sources in one element meet in the same thread, so plain adds suffice.
__device__ __forceinline__ void sources_re_point(int p, int nsrc, int istep, const int *__restrict__ ispec,
                                                 const float *__restrict__ array, const float *__restrict__ stf,
                                                 float *__restrict__ re) {
  const int slot = re_slot(p % 5, (p / 5) % 5, p / 25);
  for (int s = 0; s < nsrc; s++) {
    const float sv = stf[static_cast<size_t>(istep) * nsrc + s];
#pragma unroll
    for (int c = 0; c < 3; c++)
      re[static_cast<size_t>(ispec[s]) * RE_ELEM + c * RE_STRIDE + slot] +=
          sv * array[(static_cast<size_t>(s) * 3 + c) * 125 + p];
  }
}

__global__ __launch_bounds__(128) void sources_re_kernel(int nsrc, int istep, const int *__restrict__ ispec,
                                                         const float *__restrict__ array,
                                                         const float *__restrict__ stf, float *__restrict__ re) {
  if (threadIdx.x < 125) sources_re_point(threadIdx.x, nsrc, istep, ispec, array, stf, re);
}

// Seismogram samples of receiver r by one workgroup (its first 125 threads
// hold the element's points; a workgroup of 128 or 256 threads): Lagrange
// interpolation of u, v or a at the receiver, reduced over the workgroup.
__device__ __forceinline__ void seismogram_re_block(int r, Lattice lat, int nrec, int ntypes,
                                                    const int *__restrict__ types, int isig,
                                                    const int *__restrict__ ispec, const float *__restrict__ h,
                                                    const float *__restrict__ re, Fields f, float dto2,
                                                    float *__restrict__ seismo, int need_a, float (*red)[2]) {
  const int p = threadIdx.x;
  const int np = f.nglob_pad;
  float hw = 0.0f;
  int L = 0;
  float acc[3] = {0, 0, 0};
  if (p < 125) {
    const int e = ispec[r];
    const int ex = e % lat.nx, ey = (e / lat.nx) % lat.ny, ez = e / (lat.nx * lat.ny);
    const int i = p % 5, j = (p / 5) % 5, k = p / 25;
    L = tiled_index(lat, ex, ey, ez, i, j, k);
    hw = h[static_cast<size_t>(r) * 125 + p];
    // the assembled acceleration only for velocity and acceleration traces
    if (need_a) assemble_tiled(lat, re, ex + (i >> 2), ey + (j >> 2), ez + (k >> 2), i & 3, j & 3, k & 3, acc);
  }
  for (int it = 0; it < ntypes; it++) {
    const int type = types[it];
    float val[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float x = 0.0f;
      if (p < 125) {
        const size_t o = static_cast<size_t>(c) * np + L;
        if (type == 0) {
          x = f.u[o];
        } else {
          const float a = acc[c] * f.minv[o];
          x = type == 1 ? f.v[o] + dto2 * a : a;
        }
      }
      val[c] = hw * x;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float s = val[c];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
      // the 125 points are in the first two waves; later waves hold zeros
      if ((p & 63) == 0 && p < 128) red[c][p >> 6] = s;
    }
    __syncthreads();
    if (p == 0) {
#pragma unroll
      for (int c = 0; c < 3; c++)
        seismo[((static_cast<size_t>(isig) * ntypes + it) * nrec + r) * 3 + c] = red[c][0] + red[c][1];
    }
    __syncthreads();
  }
}

// One thread per element point on the absorbing boundary; its entries (one per
// absorbing face through the point) are summed in entry order.
__global__ __launch_bounds__(256) void stacey_re_kernel(int nseg, const int *__restrict__ seg,
                                                        const int *__restrict__ gather,
                                                        const int *__restrict__ slot,
                                                        const float *__restrict__ normal,
                                                        const float *__restrict__ weight,
                                                        const float *__restrict__ rho_vp,
                                                        const float *__restrict__ rho_vs, Fields f,
                                                        float *__restrict__ re, int src_block, int nsrc,
                                                        int istep, const int *__restrict__ src_ispec,
                                                        const float *__restrict__ src_array,
                                                        const float *__restrict__ stf, int seis_first,
                                                        Lattice lat, int nrec, int ntypes,
                                                        const int *__restrict__ types, int isig,
                                                        const int *__restrict__ rec_ispec,
                                                        const float *__restrict__ rec_h, float dto2,
                                                        float *__restrict__ seismo) {
  // after the Stacey blocks: one block adding the sources (only launched so
  // when no source element has an absorbing face: the two touch disjoint
  // entries), then one block per receiver for displacement traces, which read
  // only u
  if (seis_first >= 0 && static_cast<int>(blockIdx.x) >= seis_first) {
    __shared__ float red[3][2];
    seismogram_re_block(blockIdx.x - seis_first, lat, nrec, ntypes, types, isig, rec_ispec, rec_h, re, f, dto2,
                        seismo, 0, red);
    return;
  }
  if (static_cast<int>(blockIdx.x) == src_block) {
    if (threadIdx.x < 125) sources_re_point(threadIdx.x, nsrc, istep, src_ispec, src_array, stf, re);
    return;
  }
  const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nseg) return;
  const int q0 = seg[sidx], q1 = seg[sidx + 1];
  const int g = gather[q0];
  const int sl = slot[q0];
  const int e = sl / 125, p = sl - (sl / 125) * 125;
  const int rs = re_slot(p % 5, (p / 5) % 5, p / 25);
  const int np = f.nglob_pad;
  float v[3], t[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int d = 0; d < 3; d++) v[d] = f.v[d * np + g];
  for (int q = q0; q < q1; q++) {
    float n[3];
    float vn = 0.0f, nn = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      n[d] = normal[q * 3 + d];
      vn += v[d] * n[d];
      nn += n[d] * n[d];
    }
    const float j1 = sqrtf(nn);
    const float w = weight[q], rvp = rho_vp[q], rvs = rho_vs[q];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const float fac = ((vn * n[d] / (j1 * j1)) * (rvp - rvs)) + v[d] * rvs;
      t[d] += -1.0f * fac * j1 * w;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) re[static_cast<size_t>(e) * RE_ELEM + d * RE_STRIDE + rs] += t[d];
}

__global__ __launch_bounds__(128) void seismogram_re_kernel(Lattice lat, int nrec, int ntypes,
                                                            const int *__restrict__ types, int isig,
                                                            const int *__restrict__ ispec,
                                                            const float *__restrict__ h,
                                                            const float *__restrict__ re, Fields f, float dto2,
                                                            float *__restrict__ seismo, int need_a) {
  __shared__ float red[3][2];
  seismogram_re_block(blockIdx.x, lat, nrec, ntypes, types, isig, ispec, h, re, f, dto2, seismo, need_a, red);
}

__global__ void step_tick_kernel(int *step, int delta) {
  if (threadIdx.x == 0) *step = *step + delta;
}

__global__ void fill_kernel(float *p, size_t n, float value) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = value;
}

inline unsigned blocks(size_t n, unsigned per) { return static_cast<unsigned>((n + per - 1) / per); }

}  // namespace

void launch_stiffness_2d(hipStream_t s, int first, int count, bool atomic, const int *gather, const float *xix,
                         const float *xiz, const float *gammax, const float *gammaz, const float *jac,
                         const float *kappa, const float *mu, const GllConst &q, Fields f, int rec_mode,
                         const int *rec_index, float *rec, const int *step, long long rec_stride, const float *geo) {
  if (count <= 0) return;
  if (atomic)
    hipLaunchKernelGGL(stiffness_2d_kernel<true>, dim3(blocks(count, E2)), dim3(128), 0, s, first, count, gather, xix,
                       xiz, gammax, gammaz, jac, kappa, mu, q, f, rec_mode, rec_index, rec, step, rec_stride, geo);
  else
    hipLaunchKernelGGL(stiffness_2d_kernel<false>, dim3(blocks(count, E2)), dim3(128), 0, s, first, count, gather, xix,
                       xiz, gammax, gammaz, jac, kappa, mu, q, f, rec_mode, rec_index, rec, step, rec_stride, geo);
}

template <int E>
void launch_stiffness_2d_e(hipStream_t s, int first, int count, bool atomic, const int *gather, const float *xix,
                           const float *xiz, const float *gammax, const float *gammaz, const float *jac,
                           const float *kappa, const float *mu, const GllConst &q, Fields f) {
  if (count <= 0) return;
  if (atomic)
    hipLaunchKernelGGL((stiffness_2d_kernel<true, E>), dim3(blocks(count, E)), dim3(threads_for(E)), 0, s, first,
                       count, gather, xix, xiz, gammax, gammaz, jac, kappa, mu, q, f, 0, nullptr, nullptr, nullptr, 0LL,
                       nullptr);
  else
    hipLaunchKernelGGL((stiffness_2d_kernel<false, E>), dim3(blocks(count, E)), dim3(threads_for(E)), 0, s, first,
                       count, gather, xix, xiz, gammax, gammaz, jac, kappa, mu, q, f, 0, nullptr, nullptr, nullptr, 0LL,
                       nullptr);
}

void launch_stiffness_2d_tuned(hipStream_t s, int elems_per_block, int first, int count, bool atomic,
                               const int *gather, const float *xix, const float *xiz, const float *gammax,
                               const float *gammaz, const float *jac, const float *kappa, const float *mu,
                               const GllConst &q, Fields f) {
  switch (elems_per_block) {
    case 2:
      launch_stiffness_2d_e<2>(s, first, count, atomic, gather, xix, xiz, gammax, gammaz, jac, kappa, mu, q, f);
      break;
    case 10:
      launch_stiffness_2d_e<10>(s, first, count, atomic, gather, xix, xiz, gammax, gammaz, jac, kappa, mu, q, f);
      break;
    default:
      launch_stiffness_2d_e<5>(s, first, count, atomic, gather, xix, xiz, gammax, gammaz, jac, kappa, mu, q, f);
  }
}

void launch_stiffness_pair_2d(hipStream_t s, int first, int count, bool atomic, const int *gather,
                              const float *xix, const float *xiz, const float *gammax, const float *gammaz,
                              const float *jac, const float *kappa, const float *mu, const GllConst &q, Fields adj,
                              Fields bwd, int bwd_active, const int *rec_index, const float *rec, int frechet,
                              float dt, int nspec, float *Kmu, float *Kkappa, const int *step,
                              long long rec_stride) {
  if (count <= 0) return;
  if (atomic)
    hipLaunchKernelGGL(stiffness_pair_2d_kernel<true>, dim3(blocks(count, E2)), dim3(128), 0, s, first, count, gather,
                       xix, xiz, gammax, gammaz, jac, kappa, mu, q, adj, bwd, bwd_active, rec_index, rec, frechet, dt,
                       Kmu, Kkappa, step, rec_stride);
  else
    hipLaunchKernelGGL(stiffness_pair_2d_kernel<false>, dim3(blocks(count, E2)), dim3(128), 0, s, first, count, gather,
                       xix, xiz, gammax, gammaz, jac, kappa, mu, q, adj, bwd, bwd_active, rec_index, rec, frechet, dt,
                       Kmu, Kkappa, step, rec_stride);
}

void launch_step_tick(hipStream_t s, int *step, int delta) {
  hipLaunchKernelGGL(step_tick_kernel, dim3(1), dim3(64), 0, s, step, delta);
}

void launch_update_pair(hipStream_t s, int ncomp, Fields adj, Fields bwd, NewmarkConst nc, int bwd_correct, float *R) {
  const int n4 = adj.nglob_pad / 4;
  hipLaunchKernelGGL(update_pair_kernel, dim3(blocks(n4, 256)), dim3(256), 0, s, n4, ncomp, adj, bwd, nc, bwd_correct,
                     R);
}

void launch_rho_product(hipStream_t s, int ncomp, Fields adj, Fields bwd, float dt, float *R) {
  hipLaunchKernelGGL(rho_product_kernel, dim3(blocks(adj.nglob_pad, 256)), dim3(256), 0, s, adj.nglob_pad, ncomp, adj,
                     bwd, dt, R);
}

void launch_frechet_finish(hipStream_t s, int npts, const int *gather, const float *kappa, const float *mu,
                           const float *rho, const float *R, const float *Kmu, const float *Kkappa, float *out) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(frechet_finish_kernel, dim3(blocks(npts, 256)), dim3(256), 0, s, npts, gather, kappa, mu, rho, R,
                     Kmu, Kkappa, out);
}

void launch_frechet_2d(hipStream_t s, int nspec, const int *gather, const float *xix, const float *xiz,
                       const float *gammax, const float *gammaz, const float *kappa, const float *mu, const float *rho,
                       const GllConst &q, Fields adj, Fields bwd, float dt, float *K) {
  if (nspec <= 0) return;
  hipLaunchKernelGGL(frechet_2d_kernel, dim3(blocks(nspec, E2)), dim3(128), 0, s, nspec, gather, xix, xiz, gammax,
                     gammaz, kappa, mu, rho, q, adj, bwd, dt, K);
}

void launch_stiffness_3d(hipStream_t s, int first, int count, bool atomic, const int *gather, const float *kappa,
                         const float *mu, float xix, float ety, float gaz, float jac, const GllConst &q, Fields f) {
  if (count <= 0) return;
  if (atomic)
    hipLaunchKernelGGL(stiffness_3d_kernel<true>, dim3(blocks(count, E3)), dim3(128), 0, s, first, count, gather, kappa,
                       mu, xix, ety, gaz, jac, q, f);
  else
    hipLaunchKernelGGL(stiffness_3d_kernel<false>, dim3(blocks(count, E3)), dim3(128), 0, s, first, count, gather,
                       kappa, mu, xix, ety, gaz, jac, q, f);
}

void launch_stacey(hipStream_t s, int dim, int nbnd, const int *gather, const float *normal, const float *weight,
                   const float *rho_vp, const float *rho_vs, Fields f, const int *bnd_rec, float *rec,
                   const int *step, long long rec_stride) {
  if (nbnd <= 0) return;
  if (dim == 2)
    hipLaunchKernelGGL(stacey_kernel<2>, dim3(blocks(nbnd, 256)), dim3(256), 0, s, nbnd, gather, normal, weight,
                       rho_vp, rho_vs, f, bnd_rec, rec, step, rec_stride);
  else
    hipLaunchKernelGGL(stacey_kernel<3>, dim3(blocks(nbnd, 256)), dim3(256), 0, s, nbnd, gather, normal, weight,
                       rho_vp, rho_vs, f, bnd_rec, rec, step, rec_stride);
}

void launch_sources(hipStream_t s, int dim, int npe, int nsrc, int istep, const int *order, const int *ispec,
                    const int *gather, const float *array, const float *stf, Fields f, const int *step, int norder) {
  if (norder < 0) norder = nsrc;
  if (norder > nsrc) norder = nsrc;
  if (norder <= 0) return;
  if (dim == 2)
    hipLaunchKernelGGL(sources_kernel<2>, dim3(norder), dim3(64), 0, s, npe, nsrc, istep, order, ispec, gather, array,
                       stf, f, step, norder);
  else
    hipLaunchKernelGGL(sources_kernel<3>, dim3(norder), dim3(128), 0, s, npe, nsrc, istep, order, ispec, gather,
                       array, stf, f, step, norder);
}

void launch_seismograms(hipStream_t s, int dim, int npe, int nrec, int ntypes, const int *types, int isig,
                        const int *ispec, const int *gather, const float *h, Fields f, float dto2, int corrected,
                        float *seismo, const PointGeom2D &pg, const GllConst &q, const int *step, int nsb,
                        int sig_top, ForceSum2D fs, const int *rec_list, int nlist) {
  // nlist >= 0: the receivers rec_list[0..nlist) only (an empty list launches
  // nothing, whatever rec_list holds); -1: every receiver
  if (nlist > 0 && !rec_list) throw std::invalid_argument("receiver list missing");
  const int nb = nlist >= 0 ? nlist : nrec;
  if (nb <= 0 || ntypes <= 0) return;
  if (dim == 2)
    hipLaunchKernelGGL(seismogram_kernel<2>, dim3(nb), dim3(128), 0, s, npe, nrec, ntypes, types, isig, ispec,
                       gather, h, f, dto2, corrected, seismo, pg, q, step, nsb, sig_top, fs, rec_list);
  else
    hipLaunchKernelGGL(seismogram_kernel<3>, dim3(nb), dim3(128), 0, s, npe, nrec, ntypes, types, isig, ispec,
                       gather, h, f, dto2, corrected, seismo, pg, q, step, nsb, sig_top, ForceSum2D{}, rec_list);
}

// element groups per workgroup of elemforce_2d
constexpr int EF2_GROUPS = 2;

template <int GEOM>
static void launch_elemforce_2d_geom(hipStream_t s, int nspec, const int *gather, const Geom2D &gm,
                                     const float *kappa, const float *mu, const float *kmu, const GllConst &q,
                                     const float *u, int nglob_pad, float *re2) {
  const dim3 grid(blocks(nspec, EF2_GROUPS * E2));
  if (kmu)
    hipLaunchKernelGGL((elemforce_2d_kernel<GEOM, true, EF2_GROUPS>), grid, dim3(128), 0, s, nspec, gather, gm, kappa,
                       mu, kmu, q, u, nglob_pad, re2);
  else
    hipLaunchKernelGGL((elemforce_2d_kernel<GEOM, false, EF2_GROUPS>), grid, dim3(128), 0, s, nspec, gather, gm,
                       kappa, mu, kmu, q, u, nglob_pad, re2);
}

void launch_elemforce_2d(hipStream_t s, int nspec, const int *gather, const Geom2D &gm, const float *kappa,
                         const float *mu, const float *kmu, const GllConst &q, const float *u, int nglob_pad,
                         float *re2) {
  if (nspec <= 0) return;
  switch (gm.mode) {
    case GEOM2D_POINT:
      launch_elemforce_2d_geom<GEOM2D_POINT>(s, nspec, gather, gm, kappa, mu, kmu, q, u, nglob_pad, re2);
      break;
    case GEOM2D_RECORD:
      launch_elemforce_2d_geom<GEOM2D_RECORD>(s, nspec, gather, gm, kappa, mu, kmu, q, u, nglob_pad, re2);
      break;
    case GEOM2D_CTRL4:
      launch_elemforce_2d_geom<GEOM2D_CTRL4>(s, nspec, gather, gm, kappa, mu, kmu, q, u, nglob_pad, re2);
      break;
    case GEOM2D_CTRL9:
      launch_elemforce_2d_geom<GEOM2D_CTRL9>(s, nspec, gather, gm, kappa, mu, kmu, q, u, nglob_pad, re2);
      break;
    default:
      throw std::invalid_argument("elemforce_2d: unknown geometry form");
  }
}

void launch_stacey_re2(hipStream_t s, int nseg, const int *seg, const int *gather, const int *slot,
                       const float *normal, const float *weight, const float *rho_vp, const float *rho_vs, Fields f,
                       float *re2) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(stacey_re2_kernel, dim3(blocks(nseg, 256)), dim3(256), 0, s, nseg, seg, gather, slot, normal,
                     weight, rho_vp, rho_vs, f, re2);
}

void launch_sources_re2(hipStream_t s, int nsrc, int istep, const int *order, const int *ispec, const float *array,
                        const float *stf, float *re2, const int *step) {
  if (nsrc <= 0) return;
  hipLaunchKernelGGL(sources_re2_kernel, dim3(1), dim3(64), 0, s, nsrc, istep, order, ispec, array, stf, re2, step);
}

void launch_assemble_update_2d(hipStream_t s, int nglob, ForceSum2D fs, Fields f, NewmarkConst nc, int divide_only) {
  if (nglob <= 0) return;
  hipLaunchKernelGGL(assemble_update_2d_kernel, dim3(blocks(nglob, 256)), dim3(256), 0, s, nglob, fs, f, nc,
                     divide_only);
}

void launch_update_fused(hipStream_t s, int ncomp, Fields f, NewmarkConst nc) {
  const int n4 = f.nglob_pad / 4;
  hipLaunchKernelGGL(update_fused_kernel, dim3(blocks(n4, 256)), dim3(256), 0, s, n4, ncomp, f, nc);
}

void launch_predictor(hipStream_t s, int ncomp, Fields f, NewmarkConst nc, bool rounded) {
  if (rounded)
    hipLaunchKernelGGL(predictor_kernel<true>, dim3(blocks(f.nglob_pad, 256)), dim3(256), 0, s, f.nglob_pad, ncomp,
                       f, nc);
  else
    hipLaunchKernelGGL(predictor_kernel<false>, dim3(blocks(f.nglob_pad, 256)), dim3(256), 0, s, f.nglob_pad, ncomp,
                       f, nc);
}

void launch_divide_corrector(hipStream_t s, int ncomp, Fields f, NewmarkConst nc, int divide, int correct,
                             bool rounded) {
  if (rounded)
    hipLaunchKernelGGL(divide_corrector_kernel<true>, dim3(blocks(f.nglob_pad, 256)), dim3(256), 0, s, f.nglob_pad,
                       ncomp, f, nc, divide, correct);
  else
    hipLaunchKernelGGL(divide_corrector_kernel<false>, dim3(blocks(f.nglob_pad, 256)), dim3(256), 0, s, f.nglob_pad,
                       ncomp, f, nc, divide, correct);
}

void launch_fill(hipStream_t s, float *p, size_t n, float value) {
  if (n == 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, p, n, value);
}

Gll3 scaled_gll3(const GllConst &q, float xix, float ety, float gaz, float jac) {
  Gll3 g{};
  for (int a = 0; a < 5; a++) {
    g.w[a] = q.w[a];
    for (int l = 0; l < 5; l++) {
      g.hx[a][l] = xix * q.hp[a][l];
      g.hy[a][l] = ety * q.hp[a][l];
      g.hz[a][l] = gaz * q.hp[a][l];
      g.wx[a][l] = (jac * xix) * q.hpw[a][l];
      g.wy[a][l] = (jac * ety) * q.hpw[a][l];
      g.wz[a][l] = (jac * gaz) * q.hpw[a][l];
    }
  }
  return g;
}

void launch_elemforce_3d(hipStream_t s, Lattice lat, const float *kappa, const float *mu, const float *kmu,
                         float xix, float ety, float gaz, float jac, const GllConst &q, const float *u,
                         int nglob_pad, float *re) {
  const int nspec = lat.nx * lat.ny * lat.nz;
  if (nspec <= 0) return;
  const Gll3 g = scaled_gll3(q, xix, ety, gaz, jac);
  const double inv_nx = 1.0 / lat.nx, inv_nxy = 1.0 / (static_cast<double>(lat.nx) * lat.ny);
  if (kmu)
    hipLaunchKernelGGL(elemforce_3d_kernel<true>, dim3(blocks(nspec, E3)), dim3(128), 0, s, lat, kappa, mu, kmu, g,
                       u, nglob_pad, inv_nx, inv_nxy, re);
  else
    hipLaunchKernelGGL(elemforce_3d_kernel<false>, dim3(blocks(nspec, E3)), dim3(128), 0, s, lat, kappa, mu, kmu, g,
                       u, nglob_pad, inv_nx, inv_nxy, re);
}

void launch_sources_re(hipStream_t s, int nsrc, int istep, const int *ispec, const float *array, const float *stf,
                       float *re) {
  if (nsrc <= 0) return;
  hipLaunchKernelGGL(sources_re_kernel, dim3(1), dim3(128), 0, s, nsrc, istep, ispec, array, stf, re);
}

void launch_stacey_re(hipStream_t s, int nseg, const int *seg, const int *gather, const int *slot,
                      const float *normal, const float *weight, const float *rho_vp, const float *rho_vs, Fields f,
                      float *re) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(stacey_re_kernel, dim3(blocks(nseg, 256)), dim3(256), 0, s, nseg, seg, gather, slot, normal,
                     weight, rho_vp, rho_vs, f, re, -1, 0, 0, nullptr, nullptr, nullptr, -1, Lattice{}, 0, 0, nullptr,
                     0, nullptr, nullptr, 0.0f, nullptr);
}
void launch_stacey_sources_re(hipStream_t s, int nseg, const int *seg, const int *gather, const int *slot,
                              const float *normal, const float *weight, const float *rho_vp, const float *rho_vs,
                              Fields f, int nsrc, int istep, const int *src_ispec, const float *src_array,
                              const float *stf, float *re) {
  if (nseg <= 0) {
    launch_sources_re(s, nsrc, istep, src_ispec, src_array, stf, re);
    return;
  }
  const int nb = static_cast<int>(blocks(nseg, 256));
  const int src_block = nsrc > 0 ? nb : -1;
  hipLaunchKernelGGL(stacey_re_kernel, dim3(nb + (nsrc > 0 ? 1 : 0)), dim3(256), 0, s, nseg, seg, gather, slot,
                     normal, weight, rho_vp, rho_vs, f, re, src_block, nsrc, istep, src_ispec, src_array, stf, -1,
                     Lattice{}, 0, 0, nullptr, 0, nullptr, nullptr, 0.0f, nullptr);
}
void launch_stacey_sources_seismo_re(hipStream_t s, int nseg, const int *seg, const int *gather, const int *slot,
                                     const float *normal, const float *weight, const float *rho_vp,
                                     const float *rho_vs, Fields f, int nsrc, int istep, const int *src_ispec,
                                     const float *src_array, const float *stf, float *re, Lattice lat, int nrec,
                                     int ntypes, const int *types, int isig, const int *rec_ispec, const float *rec_h,
                                     float dto2, float *seismo) {
  const int nb = nseg > 0 ? static_cast<int>(blocks(nseg, 256)) : 0;
  const int src_block = nsrc > 0 ? nb : -1;
  const int seis_first = nb + (nsrc > 0 ? 1 : 0);
  const int nrec_blocks = nrec > 0 && ntypes > 0 ? nrec : 0;
  if (seis_first + nrec_blocks == 0) return;
  hipLaunchKernelGGL(stacey_re_kernel, dim3(seis_first + nrec_blocks), dim3(256), 0, s, nseg, seg, gather, slot,
                     normal, weight, rho_vp, rho_vs, f, re, src_block, nsrc, istep, src_ispec, src_array, stf,
                     nrec_blocks > 0 ? seis_first : -1, lat, nrec, ntypes, types, isig, rec_ispec, rec_h, dto2,
                     seismo);
}

void launch_assemble_update_3d(hipStream_t s, Lattice lat, const float *re, const float *zero, Fields f,
                               NewmarkConst nc, const float *minv1, int amask) {
  const size_t n = static_cast<size_t>(lat.TX) * lat.TY * lat.TZ * 16;
  hipLaunchKernelGGL(assemble_update_3d_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, lat, re, zero, f, nc, minv1,
                     amask);
}

void launch_assemble_3d(hipStream_t s, Lattice lat, const float *re, Fields f, int divide) {
  const size_t n = static_cast<size_t>(lat.TX) * lat.TY * lat.TZ * 64;
  hipLaunchKernelGGL(assemble_3d_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, lat, re, f, divide);
}

void launch_seismograms_re(hipStream_t s, Lattice lat, int nrec, int ntypes, const int *types, int isig,
                           const int *ispec, const float *h, const float *re, Fields f, float dto2, float *seismo,
                           bool need_a) {
  if (nrec <= 0 || ntypes <= 0) return;
  hipLaunchKernelGGL(seismogram_re_kernel, dim3(nrec), dim3(128), 0, s, lat, nrec, ntypes, types, isig, ispec, h, re, f,
                     dto2, seismo, need_a ? 1 : 0);
}

}  // namespace sfx
