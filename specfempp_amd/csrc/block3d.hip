// 3-D structured path with block assembly (SFX_SCATTER_BLOCK).
//
// A workgroup owns a block of 2 x 2 x 2 elements (9 x 9 x 9 GLL lattice
// points).  It loads the block's displacement once, computes the eight
// element forces (the stiffness of compute_stiffness_interaction.tpp:143-259
// on regular hexahedra), and sums them per block point in LDS in a fixed
// order.  A point strictly inside the block receives forces from this block
// only, so its force is complete: the kernel applies the mass division, the
// corrector of step n and the predictor of step n+1 (divide_mass_matrix.tpp:42-49,
// newmark.tpp:11-116) to it directly.  Block-surface sums go to the owner
// buffer (layout in lattice.hpp); surface_update_3d adds the (up to 8) block
// sums of every surface point, again in a fixed order, and updates it.  No
// atomics; every point's force is summed in the same order each step.
//
// The fields are in the block layout (lattice.hpp): a block's 343 interior
// points are one contiguous range per component, read and written with
// coalesced accesses; its 386 surface points live in the owner slots of up to
// 8 blocks.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "lattice.hpp"

namespace sfx {

namespace {

constexpr int BP = 729;     // lattice points of a block
constexpr int BINT = 343;   // points strictly inside a block
constexpr int BSURF = 386;  // points on a block's surface

__device__ __forceinline__ int layout_index(const Lattice &lat, int I, int J, int K) {
  return block_layout_index(lat.nx >> 1, lat.ny >> 1, lat.nint_pad, I, J, K);
}

// Owner-local coordinates of the owner's own surface point o (< 169).
__device__ __forceinline__ void own_point(int o, int &li, int &lj, int &lk) {
  if (o < 64) {
    lk = 0, lj = o >> 3, li = o & 7;
  } else {
    const int r = o - 64;
    lk = 1 + r / 15;
    const int s = r - (r / 15) * 15;
    if (s < 8)
      lj = 0, li = s;
    else
      lj = s - 7, li = 0;
  }
}

// Force at surface point (li, lj, lk) of owner (ox, oy, oz): the sums of the
// blocks holding it, region order 0..7.
__device__ __forceinline__ void owner_force(const Lattice &lat, const float *__restrict__ bs, int ox, int oy, int oz,
                                            int li, int lj, int lk, float sum[3]) {
  const int nbx = lat.nx >> 1, nby = lat.ny >> 1, nbz = lat.nz >> 1;
  const float *base = bs + ((static_cast<size_t>(oz) * (nby + 1) + oy) * (nbx + 1) + ox) * BS_ELEM;
  bool first = true;
  sum[0] = sum[1] = sum[2] = 0.0f;
#pragma unroll
  for (int d = 0; d < 8; d++) {
    const int dx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
    if ((dx && li) || (dy && lj) || (dz && lk)) continue;
    const int bx = ox - dx, by = oy - dy, bz = oz - dz;
    if (bx < 0 || by < 0 || bz < 0 || bx >= nbx || by >= nby || bz >= nbz) continue;
    const float *r = base + owner_slot(d, li, lj, lk);
#pragma unroll
    for (int c = 0; c < 3; c++) sum[c] = first ? r[c * BS_STRIDE] : sum[c] + r[c * BS_STRIDE];
    first = false;
  }
}

__global__ __launch_bounds__(256) void block_step_3d_kernel(
    Lattice lat, const float *__restrict__ kappa_, const float *__restrict__ mu_, Gll3 q, Fields f, NewmarkConst nc,
    int nsrc, int istep, const int *__restrict__ src_ispec,
    const float *__restrict__ src_array, const float *__restrict__ stf, const int *__restrict__ rec_block,
    int store_a, int divide_only, float *__restrict__ bs) {
  __shared__ float ub[3][BP];
  __shared__ float sF[8][6][125];  // F(c, x) and F(c, y) per element; then the element forces

  const int nbx = lat.nx >> 1, nby = lat.ny >> 1;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int bx = blk % nbx, by = (blk / nbx) % nby, bz = blk / (nbx * nby);
  const int t = threadIdx.x;
  const int np = f.nglob_pad;

  // thread = (element el, column (i, j)) of the force phases, points k = 0..4.
  // Threads 200..255 only load and update points; they still read the
  // quadrature rows of a valid column (t % 25), so every index stays < 5.
  const bool active = t < 200;
  const int el = active ? t / 25 : 0;
  const int col = active ? t - el * 25 : t % 25;
  const int i = col % 5, j = col / 5;
  const int dx = el & 1, dy = (el >> 1) & 1, dz = el >> 2;
  const int e = ((2 * bz + dz) * lat.ny + 2 * by + dy) * lat.nx + 2 * bx + dx;
  const int o0 = (4 * dz * 9 + 4 * dy) * 9 + 4 * dx;  // block-local offset of the element's (0, 0, 0)

  // material first, so its latency overlaps the displacement loads
  float kap[5], muv[5];
  if (active) {
    const size_t base = static_cast<size_t>(e) * 125 + col;
#pragma unroll
    for (int k = 0; k < 5; k++) {
      kap[k] = kappa_[base + k * 25];
      muv[k] = mu_[base + k * 25];
    }
  }

  // block displacement: the contiguous interior, then the surface points;
  // v and M^-1 of this thread's first interior point are fetched now so their
  // latency overlaps the force phases
  const size_t ibase = static_cast<size_t>(blk) * BINT;
  float pv[3], pm[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    pv[c] = f.v[c * np + ibase + t];
    pm[c] = f.minv[c * np + ibase + t];
  }
  for (int qi = t; qi < BINT; qi += 256) {
    const int li = qi % 7 + 1, lj = (qi / 7) % 7 + 1, lk = qi / 49 + 1;
    const int p = (lk * 9 + lj) * 9 + li;
#pragma unroll
    for (int c = 0; c < 3; c++) ub[c][p] = f.u[c * np + ibase + qi];
  }
  for (int qs = t; qs < BSURF; qs += 256) {
    int bi, bj, bk;
    block_surface_point(qs, bi, bj, bk);
    const int d = layout_index(lat, 8 * bx + bi, 8 * by + bj, 8 * bz + bk);
#pragma unroll
    for (int c = 0; c < 3; c++) ub[c][(bk * 9 + bj) * 9 + bi] = f.u[c * np + d];
  }
  __syncthreads();

  float hpi[5], hpj[5], hwi[5], hwj[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    hpi[l] = q.hx[i][l];
    hpj[l] = q.hy[j][l];
    hwi[l] = q.wx[i][l];
    hwj[l] = q.wy[j][l];
  }
  const float wi = q.w[i], wj = q.w[j];
  float u[5][3], F2[5][3], r[5][3];
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) u[k][c] = ub[c][o0 + (k * 9 + j) * 9 + i];
#pragma unroll
    for (int k = 0; k < 5; k++) {
      // strain components from the scaled contractions: d[c] = du_c / dx etc.
      float dx[3] = {0, 0, 0}, dy[3] = {0, 0, 0}, dz[3] = {0, 0, 0};
#pragma unroll
      for (int l = 0; l < 5; l++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          dx[c] += hpi[l] * ub[c][o0 + (k * 9 + j) * 9 + l];
          dy[c] += hpj[l] * ub[c][o0 + (k * 9 + l) * 9 + i];
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
      sF[el][0][p] = sxx;
      sF[el][1][p] = sxy;
      sF[el][2][p] = sxz;
      sF[el][3][p] = sxy;
      sF[el][4][p] = syy;
      sF[el][5][p] = syz;
      F2[k][0] = sxz;
      F2[k][1] = syz;
      F2[k][2] = szz;
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++) {
      float r0[3] = {0, 0, 0}, r1[3] = {0, 0, 0}, r2[3] = {0, 0, 0};
#pragma unroll
      for (int l = 0; l < 5; l++) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          r0[c] += sF[el][c][k * 25 + j * 5 + l] * hwi[l];
          r1[c] += sF[el][3 + c][k * 25 + l * 5 + i] * hwj[l];
          r2[c] += F2[l][c] * q.wz[k][l];
        }
      }
      const float wjk = wj * q.w[k], wik = wi * q.w[k], wij = wi * wj;
#pragma unroll
      for (int c = 0; c < 3; c++) r[k][c] = -(wjk * r0[c] + wik * r1[c] + wij * r2[c]);
    }
    // sources in this element (compute_source_interaction.tpp:80-95), in source order
    for (int s = 0; s < nsrc; s++) {
      if (src_ispec[s] != e) continue;
      const float sv = stf[static_cast<size_t>(istep) * nsrc + s];
#pragma unroll
      for (int k = 0; k < 5; k++)
#pragma unroll
        for (int c = 0; c < 3; c++) r[k][c] += sv * src_array[(static_cast<size_t>(s) * 3 + c) * 125 + k * 25 + col];
    }
  }
  __syncthreads();
  if (active) {
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) sF[el][c][k * 25 + col] = r[k][c];
  }
  __syncthreads();

  // per block point: sum of the elements that hold it (z, y, x order)
  auto point_sum = [&](int bi, int bj, int bk, float sum[3]) {
    bool first = true;
    sum[0] = sum[1] = sum[2] = 0.0f;
#pragma unroll
    for (int ez = 0; ez < 2; ez++) {
      const int kz = bk - 4 * ez;
      if (kz < 0 || kz > 4) continue;
#pragma unroll
      for (int ey = 0; ey < 2; ey++) {
        const int ky = bj - 4 * ey;
        if (ky < 0 || ky > 4) continue;
#pragma unroll
        for (int ex = 0; ex < 2; ex++) {
          const int kx = bi - 4 * ex;
          if (kx < 0 || kx > 4) continue;
          const int le = ez * 4 + ey * 2 + ex, lp = (kz * 5 + ky) * 5 + kx;
#pragma unroll
          for (int c = 0; c < 3; c++) sum[c] = first ? sF[le][c][lp] : sum[c] + sF[le][c][lp];
          first = false;
        }
      }
    }
  };
  auto to_owner = [&](int bi, int bj, int bk, const float sum[3]) {
    int owner, slot;
    block_point_owner(nbx, nby, bx, by, bz, bi, bj, bk, owner, slot);
    float *o = bs + static_cast<size_t>(owner) * BS_ELEM + slot;
#pragma unroll
    for (int c = 0; c < 3; c++) o[c * BS_STRIDE] = sum[c];
  };
  const bool store = divide_only || store_a == 2 || (store_a == 1 && rec_block[blk]);
  for (int qi = t; qi < BINT; qi += 256) {
    const int li = qi % 7 + 1, lj = (qi / 7) % 7 + 1, lk = qi / 49 + 1;
    const int p = (lk * 9 + lj) * 9 + li;
    float sum[3];
    point_sum(li, lj, lk, sum);
    const bool pre = qi == t;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const size_t o = static_cast<size_t>(c) * np + ibase + qi;
      const float a = sum[c] * (pre ? pm[c] : f.minv[o]);
      if (store) f.a[o] = a;
      if (divide_only) continue;
      const float v1 = (pre ? pv[c] : f.v[o]) + nc.dto2 * a;
      f.u[o] = ub[c][p] + (nc.dt * v1 + nc.dt2o2 * a);
      f.v[o] = v1 + nc.dto2 * a;
    }
  }
  for (int qs = t; qs < BSURF; qs += 256) {
    int bi, bj, bk;
    block_surface_point(qs, bi, bj, bk);
    float sum[3];
    point_sum(bi, bj, bk, sum);
    to_owner(bi, bj, bk, sum);
  }
}

// One workgroup per owner: its (up to) 169 surface points get their assembled
// force and the Newmark update.
__global__ __launch_bounds__(192) void surface_update_3d_kernel(Lattice lat, const float *__restrict__ bs, Fields f,
                                                                NewmarkConst nc, int store_a, int divide_only) {
  const int nbx = lat.nx >> 1, nby = lat.ny >> 1;
  const int O = xcd_block(blockIdx.x, gridDim.x);
  const int ox = O % (nbx + 1), oy = (O / (nbx + 1)) % (nby + 1), oz = O / ((nbx + 1) * (nby + 1));
  const int t = threadIdx.x;
  if (t >= 169) return;
  int li, lj, lk;
  own_point(t, li, lj, lk);
  const int I = 8 * ox + li, J = 8 * oy + lj, K = 8 * oz + lk;
  if (I > 4 * lat.nx || J > 4 * lat.ny || K > 4 * lat.nz) return;
  float sum[3];
  owner_force(lat, bs, ox, oy, oz, li, lj, lk, sum);
  const int d = lat.nint_pad + O * OWNER_STRIDE + t;
  const int np = f.nglob_pad;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const size_t o = static_cast<size_t>(c) * np + d;
    const float a = sum[c] * f.minv[o];
    if (store_a || divide_only) f.a[o] = a;
    if (divide_only) continue;
    const float v1 = f.v[o] + nc.dto2 * a;
    f.u[o] = f.u[o] + (nc.dt * v1 + nc.dt2o2 * a);
    f.v[o] = v1 + nc.dto2 * a;
  }
}

// Stacey traction into the owner buffer: one thread per slot; its entries
// (element points, one per absorbing face through the point) are summed in
// entry order.  bslot = owner * 386 + slot.
__global__ __launch_bounds__(256) void stacey_bs_kernel(int nseg, const int *__restrict__ seg,
                                                        const int *__restrict__ gather,
                                                        const int *__restrict__ bslot,
                                                        const float *__restrict__ normal,
                                                        const float *__restrict__ weight,
                                                        const float *__restrict__ rho_vp,
                                                        const float *__restrict__ rho_vs, Fields f,
                                                        float *__restrict__ bs) {
  const int sidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (sidx >= nseg) return;
  const int q0 = seg[sidx], q1 = seg[sidx + 1];
  const int g = gather[q0];
  const int sl = bslot[q0];
  const int owner = sl / OWNER_SLOTS, slot = sl - (sl / OWNER_SLOTS) * OWNER_SLOTS;
  const int np = f.nglob_pad;
  float v[3], t[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int d = 0; d < 3; d++) v[d] = f.v[d * np + g];
  for (int qq = q0; qq < q1; qq++) {
    float n[3];
    float vn = 0.0f, nn = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      n[d] = normal[qq * 3 + d];
      vn += v[d] * n[d];
      nn += n[d] * n[d];
    }
    const float j1 = sqrtf(nn);
    const float w = weight[qq], rvp = rho_vp[qq], rvs = rho_vs[qq];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const float fac = ((vn * n[d] / (j1 * j1)) * (rvp - rvs)) + v[d] * rvs;
      t[d] += -1.0f * fac * j1 * w;
    }
  }
  float *o = bs + static_cast<size_t>(owner) * BS_ELEM + slot;
#pragma unroll
  for (int d = 0; d < 3; d++) o[d * BS_STRIDE] += t[d];
}

// u and v at the receiver element points before the step's block kernel
// overwrites them: snap[(r * 125 + p) * 6 + {u0, u1, u2, v0, v1, v2}].
__global__ __launch_bounds__(128) void receiver_snapshot_kernel(Lattice lat, int nrec, const int *__restrict__ ispec,
                                                                Fields f, float *__restrict__ snap) {
  const int r = blockIdx.x, p = threadIdx.x;
  if (r >= nrec || p >= 125) return;
  const int e = ispec[r];
  const int ex = e % lat.nx, ey = (e / lat.nx) % lat.ny, ez = e / (lat.nx * lat.ny);
  const int d = layout_index(lat, 4 * ex + p % 5, 4 * ey + (p / 5) % 5, 4 * ez + p / 25);
  float *o = snap + (static_cast<size_t>(r) * 125 + p) * 6;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    o[c] = f.u[c * f.nglob_pad + d];
    o[3 + c] = f.v[c * f.nglob_pad + d];
  }
}

// Seismograms of the block path: the block kernel stored a = M^-1 f of the
// block-interior points of receiver blocks in f.a; block-surface forces are
// summed from the owner buffer; u and v come from the snapshot.
__global__ __launch_bounds__(128) void seismogram_block_kernel(Lattice lat, int nrec, int ntypes,
                                                               const int *__restrict__ types, int isig,
                                                               const int *__restrict__ ispec,
                                                               const float *__restrict__ h,
                                                               const float *__restrict__ bs,
                                                               const float *__restrict__ snap, Fields f, float dto2,
                                                               float *__restrict__ seismo) {
  __shared__ float red[3][2];
  const int r = blockIdx.x;
  const int p = threadIdx.x;
  const int np = f.nglob_pad;
  float hw = 0.0f, acc[3] = {0, 0, 0}, uu[3] = {0, 0, 0}, vv[3] = {0, 0, 0};
  if (p < 125) {
    const int e = ispec[r];
    const int ex = e % lat.nx, ey = (e / lat.nx) % lat.ny, ez = e / (lat.nx * lat.ny);
    const int i = p % 5, j = (p / 5) % 5, k = p / 25;
    const int I = 4 * ex + i, J = 4 * ey + j, K = 4 * ez + k;
    const int d = layout_index(lat, I, J, K);
    hw = h[static_cast<size_t>(r) * 125 + p];
    const float *s = snap + (static_cast<size_t>(r) * 125 + p) * 6;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      uu[c] = s[c];
      vv[c] = s[3 + c];
    }
    if (!(I & 7) || !(J & 7) || !(K & 7)) {
      owner_force(lat, bs, I >> 3, J >> 3, K >> 3, I & 7, J & 7, K & 7, acc);
#pragma unroll
      for (int c = 0; c < 3; c++) acc[c] *= f.minv[static_cast<size_t>(c) * np + d];
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) acc[c] = f.a[static_cast<size_t>(c) * np + d];
    }
  }
  for (int it = 0; it < ntypes; it++) {
    const int type = types[it];
    float val[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float a = acc[c];
      const float x = type == 0 ? uu[c] : (type == 1 ? vv[c] + dto2 * a : a);
      val[c] = hw * x;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float s = val[c];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
      if ((p & 63) == 0) red[c][p >> 6] = s;
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

inline unsigned nblocks_of(size_t n, unsigned per) { return static_cast<unsigned>((n + per - 1) / per); }

}  // namespace

size_t block3d_surface_floats(const Lattice &lat) {
  return static_cast<size_t>(lat.nx / 2 + 1) * (lat.ny / 2 + 1) * (lat.nz / 2 + 1) * BS_ELEM;
}

void launch_block_step_3d(hipStream_t s, Lattice lat, const float *kappa, const float *mu, float xix, float ety,
                          float gaz, float jac, const GllConst &q, Fields f, NewmarkConst nc, int nsrc, int istep,
                          const int *src_ispec, const float *src_array, const float *stf, const int *rec_block,
                          int store_a, int divide_only, float *bs) {
  const int nb = (lat.nx / 2) * (lat.ny / 2) * (lat.nz / 2);
  if (nb <= 0) return;
  const Gll3 g = scaled_gll3(q, xix, ety, gaz, jac);
  hipLaunchKernelGGL(block_step_3d_kernel, dim3(nb), dim3(256), 0, s, lat, kappa, mu, g, f, nc, nsrc, istep,
                     src_ispec, src_array, stf, rec_block, store_a, divide_only, bs);
}

void launch_surface_update_3d(hipStream_t s, Lattice lat, const float *bs, Fields f, NewmarkConst nc, int store_a,
                              int divide_only) {
  const int no = (lat.nx / 2 + 1) * (lat.ny / 2 + 1) * (lat.nz / 2 + 1);
  hipLaunchKernelGGL(surface_update_3d_kernel, dim3(no), dim3(192), 0, s, lat, bs, f, nc, store_a, divide_only);
}

void launch_stacey_bs(hipStream_t s, int nseg, const int *seg, const int *gather, const int *bslot,
                      const float *normal, const float *weight, const float *rho_vp, const float *rho_vs, Fields f,
                      float *bs) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(stacey_bs_kernel, dim3(nblocks_of(nseg, 256)), dim3(256), 0, s, nseg, seg, gather, bslot, normal,
                     weight, rho_vp, rho_vs, f, bs);
}

void launch_receiver_snapshot(hipStream_t s, Lattice lat, int nrec, const int *ispec, Fields f, float *snap) {
  if (nrec <= 0) return;
  hipLaunchKernelGGL(receiver_snapshot_kernel, dim3(nrec), dim3(128), 0, s, lat, nrec, ispec, f, snap);
}

void launch_seismograms_block(hipStream_t s, Lattice lat, int nrec, int ntypes, const int *types, int isig,
                              const int *ispec, const float *h, const float *bs, const float *snap, Fields f,
                              float dto2, float *seismo) {
  if (nrec <= 0 || ntypes <= 0) return;
  hipLaunchKernelGGL(seismogram_block_kernel, dim3(nrec), dim3(128), 0, s, lat, nrec, ntypes, types, isig, ispec, h,
                     bs, snap, f, dto2, seismo);
}

}  // namespace sfx
