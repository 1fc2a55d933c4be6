// Batched 2-D engine: many independent simulations (meshes of different
// sizes, sources, time steps and step counts) advanced together.
//
// The reference runs one simulation per process (time_marching<forward>::run,
// time_marching.tpp:9-124); a batch replaces N such runs.  All simulations are
// packed into one device arena as one disconnected mesh:
//   - simulations are sorted by step count (longest first), so the ones still
//     running at any step are a prefix of the batch;
//   - each simulation's points start on a 1024-point boundary, so a workgroup
//     of a point-wise update belongs to one simulation and reads its Newmark
//     constants (dt differs per simulation) once;
//   - Stacey entries, sources and receivers are concatenated in the same order.
// Three scatters, chosen by the autotuner or the caller:
//   - SFX_SCATTER_GATHER: one launch per time step (batch_step_kernel) over
//     the elements in simulation-major, element order; the state is
//     double-buffered so that no grid-wide synchronisation is needed;
//   - SFX_SCATTER_COLORED / _ATOMIC: a step is sources, stiffness per color
//     (elements color-major, then by simulation, so each color's running
//     elements are a prefix of the color's range), Stacey, seismograms, fused
//     update and a step-counter tick.
// Between two simulation end-steps the launches are the same every step: such
// a segment is captured once into a HIP graph and replayed.  The autotuner
// times the scatters, elements per workgroup and graph use on the batch
// itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <climits>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "../../include/specfempp_amd.h"
#include "assembly.hpp"
#include "kernels.hpp"
#include "lattice.hpp"
#include "runtime.hpp"

namespace sfx {
namespace {

constexpr int BPTS = 1024;  // points per update workgroup (256 threads x float4)

// Corrector of step n and predictor of n + 1 (divide_mass_matrix.tpp:42-49,
// newmark.tpp:11-116) with the Newmark constants of the workgroup's simulation.
__global__ __launch_bounds__(256) void batch_update_kernel(const int *__restrict__ block_sim,
                                                           const float *__restrict__ ncs, Fields f) {
  const int sim = block_sim[blockIdx.x];
  const float dt = ncs[3 * sim], dto2 = ncs[3 * sim + 1], dt2o2 = ncs[3 * sim + 2];
  const size_t i = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
  const size_t n4 = f.nglob_pad / 4;
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const size_t o = c * n4 + i;
    float4 a = reinterpret_cast<const float4 *>(f.a)[o];
    const float4 m = reinterpret_cast<const float4 *>(f.minv)[o];
    float4 u = reinterpret_cast<const float4 *>(f.u)[o];
    float4 v = reinterpret_cast<const float4 *>(f.v)[o];
#define SFX_BUPD(X)                              \
  {                                              \
    const float ac = a.X * m.X;                  \
    const float v1 = v.X + dto2 * ac;            \
    u.X = u.X + (dt * v1 + dt2o2 * ac);          \
    v.X = v1 + dto2 * ac;                        \
  }
    SFX_BUPD(x) SFX_BUPD(y) SFX_BUPD(z) SFX_BUPD(w)
#undef SFX_BUPD
    reinterpret_cast<float4 *>(f.u)[o] = u;
    reinterpret_cast<float4 *>(f.v)[o] = v;
    reinterpret_cast<float4 *>(f.a)[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
}

// Sources (compute_source_interaction.tpp:80-95): one workgroup per source,
// one thread per element point; sources of finished simulations add nothing.
__global__ __launch_bounds__(64) void batch_sources_kernel(const int *__restrict__ step, int nstep_max,
                                                           const int *__restrict__ src_sim,
                                                           const int *__restrict__ nstep_sim,
                                                           const int *__restrict__ ispec,
                                                           const int *__restrict__ gather,
                                                           const float *__restrict__ array,
                                                           const float *__restrict__ stf, Fields f) {
  const int s = blockIdx.x, p = threadIdx.x;
  const int istep = *step;
  if (p >= 25 || istep >= nstep_sim[src_sim[s]]) return;
  const float sv = stf[static_cast<size_t>(s) * nstep_max + istep];
  const int g = gather[static_cast<size_t>(ispec[s]) * 25 + p];
#pragma unroll
  for (int c = 0; c < 2; c++) unsafeAtomicAdd(&f.a[c * f.nglob_pad + g], sv * array[(s * 2 + c) * 25 + p]);
}

// Seismograms (compute_seismogram.tpp:103-135) of every receiver whose
// simulation samples this step, from the state after the (not yet applied)
// mass division and corrector: u, v + dt/2 M^-1 a, M^-1 a.
__global__ __launch_bounds__(64) void batch_seismogram_kernel(
    const int *__restrict__ step, const int *__restrict__ rec_sim, const int *__restrict__ rec_local,
    const int *__restrict__ rec_ispec, const float *__restrict__ rec_h, const int *__restrict__ nstep_sim,
    const int *__restrict__ nsb_sim, const int *__restrict__ nrec_sim, const long long *__restrict__ seis_off,
    const float *__restrict__ ncs, int ntypes, const int *__restrict__ types, const int *__restrict__ gather,
    Fields f, float *__restrict__ seismo) {
  const int r = blockIdx.x, p = threadIdx.x;
  const int istep = *step;
  const int sim = rec_sim[r];
  if (istep >= nstep_sim[sim] || istep % nsb_sim[sim] != 0) return;
  const int isig = istep / nsb_sim[sim];
  const float dto2 = ncs[3 * sim + 1];
  const int np = f.nglob_pad;
  float hw = 0.0f;
  int g = 0;
  if (p < 25) {
    hw = rec_h[static_cast<size_t>(r) * 25 + p];
    g = gather[static_cast<size_t>(rec_ispec[r]) * 25 + p];
  }
  for (int it = 0; it < ntypes; it++) {
    const int type = types[it];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      float x = 0.0f;
      if (p < 25) {
        const size_t o = static_cast<size_t>(c) * np + g;
        const float a = f.a[o] * f.minv[o];
        x = type == 0 ? f.u[o] : (type == 1 ? f.v[o] + dto2 * a : a);
      }
      float v = hw * x;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (p == 0)
        seismo[seis_off[sim] + ((static_cast<long long>(isig) * ntypes + it) * nrec_sim[sim] + rec_local[r]) * 2 + c] =
            v;
    }
  }
}

__global__ void batch_tick_kernel(int *step) {
  if (threadIdx.x == 0) *step = *step + 1;
}

// ---------------------------------------------------------------------------
// One launch per time step (scatter SFX_SCATTER_GATHER).
//
// Launch n carries every simulation from the unassembled forces of step n - 1
// to those of step n with no grid-wide synchronisation: the state is
// double-buffered by the parity of n.  Buffers A (read) hold u, v predicted
// for step n - 1 and the element forces of step n - 1 (stiffness, Stacey,
// sources; RE2_ELEM floats per element); buffers B (written) receive the same
// for step n.  Each element point
//   1. sums the step n - 1 contributions of its global point from A in a fixed
//      order (by element, then point), divides by the mass and applies
//      corrector n - 1 and predictor n to A's u, v; every element point of a
//      global point does this with the same operands, so all hold the same
//      values, and one of them (the point's first contribution) writes u, v
//      to B;
//   2. samples the step n - 1 seismograms of receivers in its element (u from
//      A, v corrected, a) when that step is a sampling step;
//   3. computes the element's weak-form force of step n from the predicted u
//      (as elemforce_2d_kernel), adds the Stacey traction of its boundary
//      entries with the predicted v (as stacey_re2_kernel) and the sources of
//      its element in application order (as sources_re2_kernel), and stores
//      the element's forces to B with coalesced 16-byte writes.
// The step index lives in two device counters: launch n reads ctr[n % 2] and
// its first workgroup writes ctr[(n + 1) % 2] = n + 1, so a captured pair of
// launches replays for any even n.  Launch nstep_max only samples.
// Each stage performs the operations of the single-engine gather path in the
// same order (elemforce_2d_kernel, stacey_re2_kernel, sources_re2_kernel,
// assemble_update_2d_kernel, seismogram_kernel's 64-lane reduction), with no
// contraction into fused multiply-adds in either, so a simulation in the
// batch gives its engine's gather-path results bit for bit.
// ---------------------------------------------------------------------------
struct StepTables {
  const int *gather;        // (nel, 25) arena point of each element point
  const float *geo;         // (nel, 8) affine geometry record, or null
  const float *xix, *xiz, *gammax, *gammaz, *jac;  // (nel, 25) when geo is null
  const float *kappa, *mu;  // (nel, 25)
  const int *elem_sim;      // (nel) simulation of each element
  const int *off;           // (npts + 1) contributions of each point: [off[g], off[g+1])
  const int *ent;           // element * RE2_ELEM + 2 p of each contribution
  const int *seg_elem;      // (nel + 1) boundary segments of each element
  const int *seg_pt;        // element point of each segment
  const int *seg_off;       // (nseg + 1) Stacey entries of each segment
  const float *bnormal, *bweight, *brvp, *brvs;
  const int *src_elem;      // (nel + 1) sources of each element, application order
  const int *src_id;
  const float *src_array;   // (nsrc, 2, 25)
  const float *stf;         // (nsrc, nstep_max)
  const int *rec_elem;      // (nel + 1) receivers of each element
  const int *rec_id;
  const float *rec_h;       // (nrec, 25)
  const int *rec_local;     // receiver index within its simulation
  const int *nstep, *nsb, *nrec_sim;
  const long long *seis_off;
  const float *ncs;         // (nsim, 3) dt, dt/2, dt^2/2
  const int *types;
  const float *minv;        // (2, npts)
  float *seismo;
  int ntypes, nstep_max, npts;
};

struct StepState {
  float *u[2], *v[2], *re[2];
  int *ctr;  // (2)
};

template <int E> constexpr int step_threads() { return (E * 25 + 63) / 64 * 64; }
// contributions of a point loaded together before the sum (the rest one by one)
constexpr int STEP_BATCH = 4;

template <int E>
__global__ __launch_bounds__(step_threads<E>()) void batch_step_kernel(int par, int nel, StepTables T, StepState S,
                                                                      GllConst q) {
  // every product and sum rounded on its own, as in the engine's gather-path
  // kernels, so both give the same bits
#pragma clang fp contract(off)
  __shared__ float su[E][2][25];
  __shared__ float sF[E][4][25];  // F(c, d) as [2c + d]
  __shared__ __align__(16) float st[E * RE2_ELEM];
  __shared__ float ss[E][3][2][25];  // samples: u, corrected v, a

  const int n = S.ctr[par];
  if (blockIdx.x == 0 && threadIdx.x == 0) S.ctr[par ^ 1] = n + 1;
  const int t = threadIdx.x;
  const int el = t / 25;
  const int p = t - el * 25;
  const int iz = p / 5, ix = p - (p / 5) * 5;
  const int first = xcd_block(blockIdx.x, gridDim.x) * E;
  const int e = first + el;
  const bool inside = (el < E) && (e < nel);
  const int np = T.npts;
  const float *UA = S.u[par], *VA = S.v[par], *reA = S.re[par];
  float *UB = S.u[par ^ 1], *VB = S.v[par ^ 1], *reB = S.re[par ^ 1];

  // Every lane issues all of its loads here, before the first barrier and
  // without branches around them (lanes past the last element re-read it; all
  // addresses are valid): the element's simulation and the point's gather
  // index, then the point's contribution bounds, state and M^-1, the
  // simulation's constants, the first STEP_BATCH contributions and the
  // element's properties and geometry, so their latencies overlap.
  const int ec = e < nel ? e : nel - 1;
  const size_t idx = static_cast<size_t>(ec) * 25 + p;
  const int sim_l = T.elem_sim[ec];
  const int g = T.gather[idx];
  const int ns = T.nstep[sim_l], nsb = T.nsb[sim_l];
  const float dt = T.ncs[3 * sim_l], dto2 = T.ncs[3 * sim_l + 1], dt2o2 = T.ncs[3 * sim_l + 2];
  const int k0 = T.off[g], k1 = T.off[g + 1];
  float minv[2], ua[2], va[2];
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const size_t o = static_cast<size_t>(c) * np + g;
    minv[c] = T.minv[o];
    ua[c] = UA[o];
    va[c] = VA[o];
  }
  int en[STEP_BATCH];
#pragma unroll
  for (int j = 0; j < STEP_BATCH; j++) en[j] = T.ent[k0 + j < k1 ? k0 + j : k1 - 1];
  float2 rj[STEP_BATCH];
#pragma unroll
  for (int j = 0; j < STEP_BATCH; j++) rj[j] = *reinterpret_cast<const float2 *>(reA + en[j]);
  const float kappa = T.kappa[idx], mu = T.mu[idx];
  const size_t gr = static_cast<size_t>(ec) * 8;
  const float xix = T.geo ? T.geo[gr] : T.xix[idx];
  const float xiz = T.geo ? T.geo[gr + 1] : T.xiz[idx];
  const float gammax = T.geo ? T.geo[gr + 2] : T.gammax[idx];
  const float gammaz = T.geo ? T.geo[gr + 3] : T.gammaz[idx];
  const float J = T.geo ? T.geo[gr + 4] : T.jac[idx];
  float hpx[5], hpz[5], hwx[5], hwz[5];
#pragma unroll
  for (int l = 0; l < 5; l++) {
    hpx[l] = q.hp[ix][l];
    hpz[l] = q.hp[iz][l];
    hwx[l] = q.hpw[ix][l];
    hwz[l] = q.hpw[iz][l];
  }
  const float wx = q.w[ix], wz = q.w[iz];

  int sim = 0;
  bool forces = false, sample = false;
  if (inside) {
    sim = sim_l;
    forces = n < ns;
    sample = n >= 1 && n - 1 < ns && (n - 1) % nsb == 0;
  }
  float un[2] = {0.0f, 0.0f}, vn[2] = {0.0f, 0.0f};
  if (forces || sample) {
    float sum[2] = {0.0f, 0.0f};
#pragma unroll
    for (int j = 0; j < STEP_BATCH; j++)
      if (k0 + j < k1) {
        sum[0] += rj[j].x;
        sum[1] += rj[j].y;
      }
    for (int k = k0 + STEP_BATCH; k < k1; k++) {
      const float2 r = *reinterpret_cast<const float2 *>(reA + T.ent[k]);
      sum[0] += r.x;
      sum[1] += r.y;
    }
    const bool owner = forces && en[0] == e * RE2_ELEM + 2 * p;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const size_t o = static_cast<size_t>(c) * np + g;
      const float a = sum[c] * minv[c];
      const float u0 = ua[c];
      const float v1 = va[c] + dto2 * a;
      un[c] = u0 + (dt * v1 + dt2o2 * a);
      vn[c] = v1 + dto2 * a;
      if (sample) {
        ss[el][0][c][p] = u0;
        ss[el][1][c][p] = v1;
        ss[el][2][c][p] = a;
      }
      if (owner) {
        UB[o] = un[c];
        VB[o] = vn[c];
      }
    }
  }
  if (inside) {
    su[el][0][p] = un[0];
    su[el][1][p] = un[1];
  }
  __syncthreads();

  // step n - 1 samples: one thread per receiver; the reduction is the 64-lane
  // shuffle tree of seismogram_kernel (values in lanes 0-24, zeros above)
  if (sample && p == 0) {
    const int isig = (n - 1) / T.nsb[sim];
    for (int k = T.rec_elem[e]; k < T.rec_elem[e + 1]; k++) {
      const int r = T.rec_id[k];
      const float *hw = T.rec_h + static_cast<size_t>(r) * 25;
      for (int it = 0; it < T.ntypes; it++) {
        const int type = T.types[it];
#pragma unroll
        for (int c = 0; c < 2; c++) {
          float s32[32];
#pragma unroll
          for (int i = 0; i < 32; i++) s32[i] = (i < 25 ? hw[i] * ss[el][type][c][i] : 0.0f) + 0.0f;
#pragma unroll
          for (int w = 16; w > 0; w >>= 1)
#pragma unroll
            for (int i = 0; i < w; i++) s32[i] = s32[i] + s32[i + w];
          T.seismo[T.seis_off[sim] + ((static_cast<long long>(isig) * T.ntypes + it) * T.nrec_sim[sim] +
                                      T.rec_local[r]) * 2 + c] = s32[0] + 0.0f;
        }
      }
    }
  }

  if (forces) {
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
    const float lp2m = kappa + (4.0f / 3.0f) * mu;
    const float lambda = lp2m - 2.0f * mu;
    const float sxx = lp2m * du[0][0] + lambda * du[1][1];
    const float szz = lp2m * du[1][1] + lambda * du[0][0];
    const float sxz = mu * (du[0][1] + du[1][0]);
    sF[el][0][p] = J * (sxx * xix + sxz * xiz);
    sF[el][1][p] = J * (sxx * gammax + sxz * gammaz);
    sF[el][2][p] = J * (sxz * xix + szz * xiz);
    sF[el][3][p] = J * (sxz * gammax + szz * gammaz);
  }
  __syncthreads();
  if (inside) {
    float r[2] = {0.0f, 0.0f};
    if (forces) {
      float t1[2] = {0.0f, 0.0f}, t2[2] = {0.0f, 0.0f};
#pragma unroll
      for (int l = 0; l < 5; l++) {
#pragma unroll
        for (int c = 0; c < 2; c++) {
          t1[c] += sF[el][2 * c][iz * 5 + l] * hwx[l];
          t2[c] += sF[el][2 * c + 1][l * 5 + ix] * hwz[l];
        }
      }
#pragma unroll
      for (int c = 0; c < 2; c++) r[c] = -(wz * t1[c] + wx * t2[c]);
      // Stacey traction with the predicted velocity (the entries of this
      // element point, summed in order, then added once)
      for (int sg = T.seg_elem[e]; sg < T.seg_elem[e + 1]; sg++) {
        if (T.seg_pt[sg] != p) continue;
        float sum[2] = {0.0f, 0.0f};
        for (int qq = T.seg_off[sg]; qq < T.seg_off[sg + 1]; qq++) {
          const float nv[2] = {T.bnormal[2 * qq], T.bnormal[2 * qq + 1]};
          const float vnn = vn[0] * nv[0] + vn[1] * nv[1];
          const float nn = nv[0] * nv[0] + nv[1] * nv[1];
          const float j1 = sqrtf(nn);
          const float w = T.bweight[qq], rvp = T.brvp[qq], rvs = T.brvs[qq];
#pragma unroll
          for (int d = 0; d < 2; d++) {
            const float fac = ((vnn * nv[d] / (j1 * j1)) * (rvp - rvs)) + vn[d] * rvs;
            sum[d] += -1.0f * fac * j1 * w;
          }
        }
        r[0] += sum[0];
        r[1] += sum[1];
      }
      for (int k = T.src_elem[e]; k < T.src_elem[e + 1]; k++) {
        const int s = T.src_id[k];
        const float sv = T.stf[static_cast<size_t>(s) * T.nstep_max + n];
#pragma unroll
        for (int c = 0; c < 2; c++) r[c] += sv * T.src_array[(static_cast<size_t>(s) * 2 + c) * 25 + p];
      }
    }
    st[el * RE2_ELEM + 2 * p] = r[0];
    st[el * RE2_ELEM + 2 * p + 1] = r[1];
  }
  __syncthreads();
  const int nb = nel - first < E ? nel - first : E;
  float4 *dst = reinterpret_cast<float4 *>(reB + static_cast<size_t>(first) * RE2_ELEM);
  const float4 *from = reinterpret_cast<const float4 *>(st);
  for (int qv = t; qv < nb * (RE2_ELEM / 4); qv += step_threads<E>()) dst[qv] = from[qv];
}

}  // namespace
}  // namespace sfx

using namespace sfx;

struct sfx_batch {
  int nsim = 0;
  std::vector<int> order;  // batch position -> simulation index given at creation
  std::vector<int> pos_of; // simulation index -> batch position
  hipStream_t stream = nullptr;
  GllConst q{};
  int ntypes = 0;
  std::vector<int> types;
  // per batch position
  std::vector<int> nstep, nsb, nrec, nsig, nglob;
  std::vector<long long> seis_off;
  int nstep_max = 0;
  // launch ranges for k running simulations (k = 0..nsim)
  int ncolors = 0;
  std::vector<int> color_off;             // (ncolors + 1)
  std::vector<std::vector<int>> color_cnt;  // [c][k] elements of color c in the first k simulations
  std::vector<int> bnd_cnt, blk_cnt, pt_off;
  size_t npts = 0;  // padded points of the arena
  // device arena
  DevBuf<int> gather, bnd_gather, src_sim, src_ispec, rec_sim, rec_local, rec_ispec, d_types, block_sim, d_nstep, d_nsb,
      d_nrec, step;
  DevBuf<long long> d_seis_off;
  DevBuf<float> xix, xiz, gammax, gammaz, jac, kappa, mu, bnd_normal, bnd_weight, bnd_rvp, bnd_rvs, src_array, stf,
      rec_h, seismo, fields, ncs;
  int nsrc = 0, nrec_total = 0;
  // one-launch-per-step layout (SFX_SCATTER_GATHER): elements in simulation-
  // major, element order; elem_cnt_f[k] elements belong to the first k
  // simulations
  int nel_f = 0;
  std::vector<int> elem_cnt_f;
  DevBuf<int> f_gather, f_elem_sim, f_off, f_ent, f_seg_elem, f_seg_pt, f_seg_off, f_src_elem, f_src_id, f_rec_elem,
      f_rec_id, ctr;
  DevBuf<float> f_geo, f_xix, f_xiz, f_gx, f_gz, f_jac, f_kappa, f_mu, f_bn, f_bw, f_bp, f_bs, re;
  // tuning
  int scatter = SFX_SCATTER_GATHER;
  int elems_per_block = 5;
  bool use_graph = true;
  double last_run_ms = 0.0;
  std::vector<hipGraphExec_t> graph_cache;  // per k (nullptr until captured)

  // fields: u, v, a, M^-1 (2 npts each), then the second u, v of the
  // double-buffered one-launch-per-step state
  Fields f() const {
    const size_t stride = 2 * npts;
    return Fields{fields.p, fields.p + stride, fields.p + 2 * stride, fields.p + 3 * stride, static_cast<int>(npts)};
  }
  StepState step_state() const {
    const size_t stride = 2 * npts;
    const size_t nre = static_cast<size_t>(nel_f) * RE2_ELEM;
    return StepState{{fields.p, fields.p + 4 * stride},
                     {fields.p + stride, fields.p + 5 * stride},
                     {re.p, re.p + nre},
                     ctr.p};
  }
  StepTables step_tables() const {
    StepTables T{};
    T.gather = f_gather.p;
    T.geo = f_geo.p;
    T.xix = f_xix.p;
    T.xiz = f_xiz.p;
    T.gammax = f_gx.p;
    T.gammaz = f_gz.p;
    T.jac = f_jac.p;
    T.kappa = f_kappa.p;
    T.mu = f_mu.p;
    T.elem_sim = f_elem_sim.p;
    T.off = f_off.p;
    T.ent = f_ent.p;
    T.seg_elem = f_seg_elem.p;
    T.seg_pt = f_seg_pt.p;
    T.seg_off = f_seg_off.p;
    T.bnormal = f_bn.p;
    T.bweight = f_bw.p;
    T.brvp = f_bp.p;
    T.brvs = f_bs.p;
    T.src_elem = f_src_elem.p;
    T.src_id = f_src_id.p;
    T.src_array = src_array.p;
    T.stf = stf.p;
    T.rec_elem = f_rec_elem.p;
    T.rec_id = f_rec_id.p;
    T.rec_h = rec_h.p;
    T.rec_local = rec_local.p;
    T.nstep = d_nstep.p;
    T.nsb = d_nsb.p;
    T.nrec_sim = d_nrec.p;
    T.seis_off = d_seis_off.p;
    T.ncs = ncs.p;
    T.types = d_types.p;
    T.minv = fields.p + 6 * npts;
    T.seismo = seismo.p;
    T.ntypes = ntypes;
    T.nstep_max = nstep_max;
    T.npts = static_cast<int>(npts);
    return T;
  }

  ~sfx_batch() {
    clear_graphs();
    if (stream) (void)hipStreamDestroy(stream);
  }
  void clear_graphs() {
    for (auto &g : graph_cache)
      if (g) (void)hipGraphExecDestroy(g);
    graph_cache.assign(nsim + 1, nullptr);
  }

  void build(const std::vector<const sfx_engine *> &engines) {
    nsim = static_cast<int>(engines.size());
    if (nsim <= 0) throw std::invalid_argument("a batch needs at least one engine");
    for (auto *e : engines) {
      if (!e) throw std::invalid_argument("null engine");
      if (engine_host(e).dim != 2) throw std::invalid_argument("batches are implemented for 2-D engines");
      if (engine_host(e).coupled()) throw std::invalid_argument("batches are implemented for elastic meshes");
    }
    const sfx_run_desc &r0 = engine_run(engines[0]);
    ntypes = r0.nseismo_types;
    types.assign(r0.seismo_types, r0.seismo_types + ntypes);
    for (auto *e : engines) {
      const sfx_run_desc &r = engine_run(e);
      if (r.nseismo_types != ntypes || !std::equal(types.begin(), types.end(), r.seismo_types))
        throw std::invalid_argument("all engines of a batch must record the same seismogram types");
      for (int t : types)
        if (t == SFX_PRESSURE) throw std::invalid_argument("batches record displacement, velocity or acceleration");
    }
    order.resize(nsim);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return engine_run(engines[a]).nstep > engine_run(engines[b]).nstep; });
    pos_of.assign(nsim, 0);
    for (int k = 0; k < nsim; k++) pos_of[order[k]] = k;
    std::vector<const HostAssembly *> H(nsim);
    std::vector<const sfx_run_desc *> R(nsim);
    // the colored scatter needs a valid coloring of every mesh, whatever
    // scatter its engine was built with (ATOMIC and GATHER engines keep one
    // "color" in compute order)
    std::vector<ElementColoring> col(nsim);
    for (int k = 0; k < nsim; k++) {
      H[k] = &engine_host(engines[order[k]]);
      R[k] = &engine_run(engines[order[k]]);
      col[k] = element_coloring(*H[k], SFX_SCATTER_COLORED);
      // every element must sit in exactly one color slot (the packing below
      // maps elements through the coloring)
      const auto &co = col[k].color_offsets;
      if (co.size() < 2 || co.front() != 0 || co.back() != H[k]->nspec ||
          static_cast<int>(col[k].perm.size()) != H[k]->nspec)
        throw std::logic_error("batch: an element coloring is incomplete");
    }
    // quadrature constants (the same GLL rule for every engine)
    for (int i = 0; i < 5; i++) {
      q.w[i] = H[0]->w[i];
      for (int l = 0; l < 5; l++) {
        q.hp[i][l] = H[0]->hprime[i][l];
        q.hpw[i][l] = H[0]->hprime[l][i] * H[0]->w[l];
      }
    }
    // points
    pt_off.assign(nsim + 1, 0);
    nstep.resize(nsim);
    nsb.resize(nsim);
    nrec.resize(nsim);
    nsig.resize(nsim);
    nglob.resize(nsim);
    seis_off.assign(nsim + 1, 0);
    std::vector<float> ncs_h(3 * static_cast<size_t>(nsim));
    for (int k = 0; k < nsim; k++) {
      const long long next = static_cast<long long>(pt_off[k]) + (H[k]->nglob + BPTS - 1) / BPTS * BPTS;
      if (next > INT_MAX / 8) throw std::invalid_argument("batch too large");
      pt_off[k + 1] = static_cast<int>(next);
      nglob[k] = H[k]->nglob;
      nstep[k] = R[k]->nstep;
      nsb[k] = R[k]->nstep_between_samples;
      nrec[k] = H[k]->nrec;
      nsig[k] = (nstep[k] + nsb[k] - 1) / nsb[k];
      seis_off[k + 1] = seis_off[k] + static_cast<long long>(nsig[k]) * ntypes * nrec[k] * 2;
      const float dt = static_cast<float>(R[k]->dt);
      ncs_h[3 * k] = dt;
      ncs_h[3 * k + 1] = dt / 2.0;
      ncs_h[3 * k + 2] = dt * dt / 2.0;
    }
    nstep_max = nstep[0];
    npts = pt_off[nsim];
    blk_cnt.resize(nsim + 1);
    for (int k = 0; k <= nsim; k++) blk_cnt[k] = pt_off[k] / BPTS;
    std::vector<int> block_sim_h(npts / BPTS);
    for (int k = 0; k < nsim; k++)
      for (int b = pt_off[k] / BPTS; b < pt_off[k + 1] / BPTS; b++) block_sim_h[b] = k;
    // elements: color-major, then simulation, each simulation's slot order
    ncolors = 0;
    for (int k = 0; k < nsim; k++) ncolors = std::max(ncolors, static_cast<int>(col[k].color_offsets.size()) - 1);
    color_off.assign(ncolors + 1, 0);
    color_cnt.assign(ncolors, std::vector<int>(nsim + 1, 0));
    std::vector<std::vector<int>> bidx(nsim);  // compute element -> batch element, per simulation
    for (int k = 0; k < nsim; k++) bidx[k].assign(H[k]->nspec, -1);
    std::vector<int> elem_sim, elem_ce;
    for (int c = 0; c < ncolors; c++) {
      color_off[c] = static_cast<int>(elem_sim.size());
      for (int k = 0; k < nsim; k++) {
        const auto &co = col[k].color_offsets;
        if (c + 1 < static_cast<int>(co.size()))
          for (int slot = co[c]; slot < co[c + 1]; slot++) {
            bidx[k][col[k].perm[slot]] = static_cast<int>(elem_sim.size());
            elem_sim.push_back(k);
            elem_ce.push_back(col[k].perm[slot]);
          }
        color_cnt[c][k + 1] = static_cast<int>(elem_sim.size()) - color_off[c];
      }
    }
    color_off[ncolors] = static_cast<int>(elem_sim.size());
    const size_t nel = elem_sim.size();
    for (int k = 0; k < nsim; k++)
      for (int b : bidx[k])
        if (b < 0) throw std::logic_error("batch: an element was not packed");
    std::vector<int> g_h(nel * 25);
    std::vector<float> xix_h(nel * 25), xiz_h(nel * 25), gx_h(nel * 25), gz_h(nel * 25), j_h(nel * 25), k_h(nel * 25),
        m_h(nel * 25);
    for (size_t b = 0; b < nel; b++) {
      const HostAssembly &h = *H[elem_sim[b]];
      const size_t src = static_cast<size_t>(elem_ce[b]) * 25;
      for (int p = 0; p < 25; p++) {
        g_h[b * 25 + p] = h.gather[src + p] + pt_off[elem_sim[b]];
        xix_h[b * 25 + p] = h.xix[src + p];
        xiz_h[b * 25 + p] = h.xiz[src + p];
        gx_h[b * 25 + p] = h.gammax[src + p];
        gz_h[b * 25 + p] = h.gammaz[src + p];
        j_h[b * 25 + p] = h.jac[src + p];
        k_h[b * 25 + p] = h.kappa[src + p];
        m_h[b * 25 + p] = h.mu[src + p];
      }
    }
    // Stacey entries, sources, receivers, mass inverse
    std::vector<int> bg_h;
    std::vector<float> bn_h, bw_h, bp_h, bs_h;
    bnd_cnt.assign(nsim + 1, 0);
    std::vector<int> ssim_h, sisp_h;
    std::vector<float> sarr_h, stf_h;
    std::vector<int> rsim_h, rloc_h, risp_h;
    std::vector<float> rh_h;
    std::vector<float> minv_h(2 * npts, 0.0f);
    for (int k = 0; k < nsim; k++) {
      const HostAssembly &h = *H[k];
      for (size_t qq = 0; qq < h.bnd_gather.size(); qq++) {
        bg_h.push_back(h.bnd_gather[qq] + pt_off[k]);
        bn_h.push_back(h.bnd_normal[2 * qq]);
        bn_h.push_back(h.bnd_normal[2 * qq + 1]);
        bw_h.push_back(h.bnd_weight[qq]);
        bp_h.push_back(h.bnd_rho_vp[qq]);
        bs_h.push_back(h.bnd_rho_vs[qq]);
      }
      bnd_cnt[k + 1] = static_cast<int>(bg_h.size());
      for (int s : h.src_order) {
        ssim_h.push_back(k);
        sisp_h.push_back(bidx[k][h.src_ispec[s]]);
        for (int i = 0; i < 50; i++) sarr_h.push_back(h.src_array[static_cast<size_t>(s) * 50 + i]);
        for (int i = 0; i < nstep_max; i++)
          stf_h.push_back(i < nstep[k] ? h.stf[static_cast<size_t>(i) * h.nsrc + s] : 0.0f);
      }
      for (int r = 0; r < h.nrec; r++) {
        rsim_h.push_back(k);
        rloc_h.push_back(r);
        risp_h.push_back(bidx[k][h.rec_ispec[r]]);
        for (int p = 0; p < 25; p++) rh_h.push_back(h.rec_h[static_cast<size_t>(r) * 25 + p]);
      }
      for (int c = 0; c < 2; c++)
        for (int g = 0; g < h.nglob; g++)
          minv_h[c * npts + pt_off[k] + g] = h.mass_inverse[static_cast<size_t>(c) * h.nglob_pad + g];
    }
    nsrc = static_cast<int>(ssim_h.size());
    nrec_total = static_cast<int>(rsim_h.size());
    // one-launch-per-step layout: simulation k's element ce is element
    // ebase[k] + ce; a point's contributions are listed by element, then by
    // point, the order in which an engine's gather path sums them
    std::vector<int> ebase(nsim + 1, 0);
    for (int k = 0; k < nsim; k++) ebase[k + 1] = ebase[k] + H[k]->nspec;
    nel_f = ebase[nsim];
    if (static_cast<long long>(nel_f) * RE2_ELEM > INT_MAX) throw std::invalid_argument("batch too large");
    elem_cnt_f = ebase;
    bool affine = true;
    for (int k = 0; k < nsim; k++) affine = affine && !H[k]->geo2d.empty();
    std::vector<int> fg_h(static_cast<size_t>(nel_f) * 25), fsim_h(nel_f);
    std::vector<float> fk_h(fg_h.size()), fm_h(fg_h.size()), fgeo_h, fx_h, fxz_h, fgx_h, fgz_h, fj_h;
    if (affine)
      fgeo_h.resize(static_cast<size_t>(nel_f) * 8);
    else
      for (auto *v : {&fx_h, &fxz_h, &fgx_h, &fgz_h, &fj_h}) v->resize(fg_h.size());
    std::vector<int> off_h(npts + 1, 0), ent_h;
    std::vector<int> sege_h(nel_f + 1, 0), segp_h, sego_h(1, 0);
    std::vector<float> fbn_h, fbw_h, fbp_h, fbs_h;
    std::vector<int> srce_h(nel_f + 1, 0), srcid_h, rece_h(nel_f + 1, 0), recid_h;
    {
      int src_base = 0, rec_base = 0;
      std::vector<std::vector<int>> src_of(nel_f), rec_of(nel_f);
      for (int k = 0; k < nsim; k++) {
        const HostAssembly &h = *H[k];
        for (int ce = 0; ce < h.nspec; ce++) {
          const int e = ebase[k] + ce;
          fsim_h[e] = k;
          for (int p = 0; p < 25; p++) {
            const size_t d = static_cast<size_t>(e) * 25 + p, src = static_cast<size_t>(ce) * 25 + p;
            fg_h[d] = h.gather[src] + pt_off[k];
            fk_h[d] = h.kappa[src];
            fm_h[d] = h.mu[src];
            if (!affine) {
              fx_h[d] = h.xix[src];
              fxz_h[d] = h.xiz[src];
              fgx_h[d] = h.gammax[src];
              fgz_h[d] = h.gammaz[src];
              fj_h[d] = h.jac[src];
            }
          }
          if (affine) std::memcpy(&fgeo_h[static_cast<size_t>(e) * 8], &h.geo2d[static_cast<size_t>(ce) * 8], 32);
        }
        for (size_t i = 0; i < static_cast<size_t>(h.nspec) * 25; i++) off_h[pt_off[k] + h.gather[i] + 1]++;
        // Stacey entries grouped by element point (stable: edge order kept)
        std::vector<size_t> ord(h.bnd_gather.size());
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return h.bnd_slot[a] < h.bnd_slot[b]; });
        for (size_t i = 0; i < ord.size(); i++) {
          const size_t qq = ord[i];
          if (i == 0 || h.bnd_slot[qq] != h.bnd_slot[ord[i - 1]]) {
            if (i > 0) sego_h.push_back(static_cast<int>(fbw_h.size()));
            segp_h.push_back(h.bnd_slot[qq] % 25);
            sege_h[ebase[k] + h.bnd_slot[qq] / 25 + 1]++;
          }
          fbn_h.push_back(h.bnd_normal[2 * qq]);
          fbn_h.push_back(h.bnd_normal[2 * qq + 1]);
          fbw_h.push_back(h.bnd_weight[qq]);
          fbp_h.push_back(h.bnd_rho_vp[qq]);
          fbs_h.push_back(h.bnd_rho_vs[qq]);
        }
        if (!ord.empty()) sego_h.push_back(static_cast<int>(fbw_h.size()));
        for (size_t i = 0; i < h.src_order.size(); i++)
          src_of[ebase[k] + h.src_ispec[h.src_order[i]]].push_back(src_base + static_cast<int>(i));
        src_base += static_cast<int>(h.src_order.size());
        for (int r = 0; r < h.nrec; r++) rec_of[ebase[k] + h.rec_ispec[r]].push_back(rec_base + r);
        rec_base += h.nrec;
      }
      for (size_t g = 0; g < npts; g++) off_h[g + 1] += off_h[g];
      ent_h.resize(off_h[npts]);
      std::vector<int> fill(off_h.begin(), off_h.end() - 1);
      for (int k = 0; k < nsim; k++) {
        const HostAssembly &h = *H[k];
        for (int ce = 0; ce < h.nspec; ce++) {
          for (int p = 0; p < 25; p++) {
            const int g = pt_off[k] + h.gather[static_cast<size_t>(ce) * 25 + p];
            ent_h[fill[g]++] = (ebase[k] + ce) * RE2_ELEM + 2 * p;
          }
        }
      }
      for (int e = 0; e < nel_f; e++) {
        sege_h[e + 1] += sege_h[e];
        srce_h[e + 1] = srce_h[e] + static_cast<int>(src_of[e].size());
        rece_h[e + 1] = rece_h[e] + static_cast<int>(rec_of[e].size());
        srcid_h.insert(srcid_h.end(), src_of[e].begin(), src_of[e].end());
        recid_h.insert(recid_h.end(), rec_of[e].begin(), rec_of[e].end());
      }
      if (sege_h[nel_f] != static_cast<int>(segp_h.size()) || sego_h.size() != segp_h.size() + 1)
        throw std::logic_error("batch: Stacey segments are inconsistent");
    }
    // upload
    SFX_HIP(hipStreamCreate(&stream));
    hipStream_t s = stream;
    gather.upload(g_h, s);
    xix.upload(xix_h, s);
    xiz.upload(xiz_h, s);
    gammax.upload(gx_h, s);
    gammaz.upload(gz_h, s);
    jac.upload(j_h, s);
    kappa.upload(k_h, s);
    mu.upload(m_h, s);
    bnd_gather.upload(bg_h, s);
    bnd_normal.upload(bn_h, s);
    bnd_weight.upload(bw_h, s);
    bnd_rvp.upload(bp_h, s);
    bnd_rvs.upload(bs_h, s);
    src_sim.upload(ssim_h, s);
    src_ispec.upload(sisp_h, s);
    src_array.upload(sarr_h, s);
    stf.upload(stf_h, s);
    rec_sim.upload(rsim_h, s);
    rec_local.upload(rloc_h, s);
    rec_ispec.upload(risp_h, s);
    rec_h.upload(rh_h, s);
    d_types.upload(types, s);
    block_sim.upload(block_sim_h, s);
    d_nstep.upload(nstep, s);
    d_nsb.upload(nsb, s);
    d_nrec.upload(nrec, s);
    d_seis_off.upload(seis_off, s);
    ncs.upload(ncs_h, s);
    f_gather.upload(fg_h, s);
    f_elem_sim.upload(fsim_h, s);
    f_kappa.upload(fk_h, s);
    f_mu.upload(fm_h, s);
    if (affine) {
      f_geo.upload(fgeo_h, s);
    } else {
      f_xix.upload(fx_h, s);
      f_xiz.upload(fxz_h, s);
      f_gx.upload(fgx_h, s);
      f_gz.upload(fgz_h, s);
      f_jac.upload(fj_h, s);
    }
    f_off.upload(off_h, s);
    f_ent.upload(ent_h, s);
    f_seg_elem.upload(sege_h, s);
    f_seg_pt.upload(segp_h, s);
    f_seg_off.upload(sego_h, s);
    f_bn.upload(fbn_h, s);
    f_bw.upload(fbw_h, s);
    f_bp.upload(fbp_h, s);
    f_bs.upload(fbs_h, s);
    f_src_elem.upload(srce_h, s);
    f_src_id.upload(srcid_h, s);
    f_rec_elem.upload(rece_h, s);
    f_rec_id.upload(recid_h, s);
    re.alloc(2 * static_cast<size_t>(nel_f) * RE2_ELEM);
    ctr.alloc(2);
    seismo.alloc(std::max<long long>(seis_off[nsim], 1));
    fields.alloc(12 * npts);
    SFX_HIP(hipMemcpyAsync(fields.p + 6 * npts, minv_h.data(), 2 * npts * sizeof(float), hipMemcpyHostToDevice, s));
    step.alloc(1);
    SFX_HIP(hipStreamSynchronize(s));
    graph_cache.assign(nsim + 1, nullptr);
  }

  // Number of simulations still running at step istep.
  int running(int istep) const {
    int k = 0;
    while (k < nsim && nstep[k] > istep) k++;
    return k;
  }
  // Simulations touched by launch n of the one-launch-per-step scatter (those
  // with step n to compute or step n - 1 to sample).
  int touched(int n) const {
    int k = 0;
    while (k < nsim && nstep[k] >= n) k++;
    return k;
  }

  // The launches of one step with k simulations running.
  void enqueue_step(int k) {
    hipStream_t s = stream;
    Fields ff = f();
    const bool atomic = scatter == SFX_SCATTER_ATOMIC;
    if (nsrc > 0)
      hipLaunchKernelGGL(batch_sources_kernel, dim3(nsrc), dim3(64), 0, s, step.p, nstep_max, src_sim.p, d_nstep.p,
                         src_ispec.p, gather.p, src_array.p, stf.p, ff);
    for (int c = 0; c < ncolors; c++)
      launch_stiffness_2d_tuned(s, elems_per_block, color_off[c], color_cnt[c][k], atomic, gather.p, xix.p, xiz.p,
                                gammax.p, gammaz.p, jac.p, kappa.p, mu.p, q, ff);
    launch_stacey(s, 2, bnd_cnt[k], bnd_gather.p, bnd_normal.p, bnd_weight.p, bnd_rvp.p, bnd_rvs.p, ff);
    if (nrec_total > 0 && ntypes > 0)
      hipLaunchKernelGGL(batch_seismogram_kernel, dim3(nrec_total), dim3(64), 0, s, step.p, rec_sim.p, rec_local.p,
                         rec_ispec.p, rec_h.p, d_nstep.p, d_nsb.p, d_nrec.p, d_seis_off.p, ncs.p, ntypes, d_types.p,
                         gather.p, ff, seismo.p);
    if (blk_cnt[k] > 0)
      hipLaunchKernelGGL(batch_update_kernel, dim3(blk_cnt[k]), dim3(256), 0, s, block_sim.p, ncs.p, ff);
    hipLaunchKernelGGL(batch_tick_kernel, dim3(1), dim3(64), 0, s, step.p);
  }

  // Launch n (n % 2 == par) of the one-launch-per-step scatter over the
  // elements of the first k simulations.
  void enqueue_launch(int par, int k) {
    const int nel = elem_cnt_f[k];
    if (nel <= 0) return;
    const StepTables T = step_tables();
    const StepState S = step_state();
    const unsigned nb = static_cast<unsigned>((nel + elems_per_block - 1) / elems_per_block);
    switch (elems_per_block) {
      case 2:
        hipLaunchKernelGGL(batch_step_kernel<2>, dim3(nb), dim3(step_threads<2>()), 0, stream, par, nel, T, S, q);
        break;
      case 10:
        hipLaunchKernelGGL(batch_step_kernel<10>, dim3(nb), dim3(step_threads<10>()), 0, stream, par, nel, T, S, q);
        break;
      default:
        hipLaunchKernelGGL(batch_step_kernel<5>, dim3(nb), dim3(step_threads<5>()), 0, stream, par, nel, T, S, q);
    }
  }

  void reset() {
    launch_fill(stream, fields.p, 6 * npts, 0.0f);
    launch_fill(stream, fields.p + 8 * npts, 4 * npts, 0.0f);
    launch_fill(stream, re.p, re.n, 0.0f);
    launch_fill(stream, seismo.p, seismo.n, 0.0f);
    const int zero[2] = {0, 0};
    SFX_HIP(hipMemcpyAsync(step.p, zero, sizeof(int), hipMemcpyHostToDevice, stream));
    SFX_HIP(hipMemcpyAsync(ctr.p, zero, 2 * sizeof(int), hipMemcpyHostToDevice, stream));
    SFX_HIP(hipStreamSynchronize(stream));
  }

  // Capture fn() into a graph executable.
  template <class Fn> hipGraphExec_t capture(Fn fn) {
    hipGraph_t g;
    SFX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      fn();
    } catch (...) {
      (void)hipStreamEndCapture(stream, &g);
      throw;
    }
    SFX_HIP(hipStreamEndCapture(stream, &g));
    hipGraphExec_t ex;
    const hipError_t err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    SFX_HIP(err);
    return ex;
  }

  // Steps [first, last) (first = 0 after reset()); with the one-launch-per-
  // step scatter, launches [first, last) (the last run needs nstep_max + 1).
  void advance(int first, int last) {
    Throttle throttle(stream);
    const bool one = scatter == SFX_SCATTER_GATHER;
    int n = first;
    while (n < last) {
      const int k = one ? touched(n) : running(n);
      int end = n;
      while (end < last && (one ? touched(end) : running(end)) == k) end++;
      if (k == 0) break;
      if (one) {
        if (use_graph) {
          // a graph holds the pair of launches (even n, odd n)
          if (n % 2 == 1) {
            enqueue_launch(1, k);
            throttle.tick();
            n++;
          }
          if (n + 1 < end && !graph_cache[k]) graph_cache[k] = capture([&] {
            enqueue_launch(0, k);
            enqueue_launch(1, k);
          });
          for (; n + 1 < end; n += 2) {
            SFX_HIP(hipGraphLaunch(graph_cache[k], stream));
            throttle.tick();
            throttle.tick();
          }
          if (n < end) {
            enqueue_launch(0, k);
            throttle.tick();
          }
        } else {
          for (int i = n; i < end; i++) {
            enqueue_launch(i % 2, k);
            throttle.tick();
          }
        }
      } else if (use_graph) {
        if (!graph_cache[k]) graph_cache[k] = capture([&] { enqueue_step(k); });
        for (int i = n; i < end; i++) {
          SFX_HIP(hipGraphLaunch(graph_cache[k], stream));
          throttle.tick();
        }
      } else {
        for (int i = n; i < end; i++) {
          enqueue_step(k);
          throttle.tick();
        }
      }
      n = end;
    }
    SFX_HIP(hipGetLastError());
  }

  // Launches (or steps) of a whole run.
  int run_length() const { return scatter == SFX_SCATTER_GATHER ? nstep_max + 1 : nstep_max; }
  // Launches of one step: 1 for the one-launch-per-step scatter; otherwise
  // sources, one per color, Stacey, seismograms, update and tick.
  int launches_per_step() const {
    return scatter == SFX_SCATTER_GATHER ? 1 : (nsrc > 0) + ncolors + 1 + (nrec_total > 0 && ntypes > 0) + 2;
  }

  double timed_steps(int nsteps) {
    hipEvent_t a, b;
    SFX_HIP(hipEventCreate(&a));
    SFX_HIP(hipEventCreate(&b));
    reset();
    advance(0, 2);  // warm-up (graph capture, first-touch)
    SFX_HIP(hipEventRecord(a, stream));
    advance(2, 2 + nsteps);
    SFX_HIP(hipEventRecord(b, stream));
    SFX_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    SFX_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
  }
};

extern "C" {

int sfx_batch_create(int n, sfx_engine *const *engines, sfx_batch **out) {
  return guarded([&] {
    if (!out || (n > 0 && !engines)) throw std::invalid_argument("null argument");
    *out = nullptr;
    auto b = std::make_unique<sfx_batch>();
    b->build(std::vector<const sfx_engine *>(engines, engines + n));
    *out = b.release();
  });
}

void sfx_batch_destroy(sfx_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  delete b;
}

int sfx_batch_tune(sfx_batch *b, int steps_per_candidate, int *choice4) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    if (steps_per_candidate < 1) throw std::invalid_argument("steps_per_candidate must be >= 1");
    const int nsteps = std::min(steps_per_candidate, std::max(b->nstep[b->nsim - 1] - 2, 1));
    double best = 1e300;
    int best_sc = SFX_SCATTER_GATHER, best_e = 5;
    bool best_g = true;
    for (int sc : {SFX_SCATTER_GATHER, SFX_SCATTER_COLORED, SFX_SCATTER_ATOMIC})
      for (int e : {2, 5, 10})
        for (bool g : {false, true}) {
          b->scatter = sc;
          b->elems_per_block = e;
          b->use_graph = g;
          b->clear_graphs();
          const double ms = b->timed_steps(nsteps);
          if (ms < best) {
            best = ms;
            best_sc = sc;
            best_e = e;
            best_g = g;
          }
        }
    b->scatter = best_sc;
    b->elems_per_block = best_e;
    b->use_graph = best_g;
    b->clear_graphs();
    b->reset();
    if (choice4) {
      choice4[0] = best_e;
      choice4[1] = best_g ? 1 : 0;
      choice4[2] = b->ncolors;
      choice4[3] = best_sc;
    }
  }, b ? b->stream : nullptr);
}

int sfx_batch_set_config(sfx_batch *b, int scatter, int elems_per_block, int use_graph) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    if (scatter != SFX_SCATTER_GATHER && scatter != SFX_SCATTER_COLORED && scatter != SFX_SCATTER_ATOMIC)
      throw std::invalid_argument("batch scatter must be SFX_SCATTER_GATHER, _COLORED or _ATOMIC");
    if (elems_per_block != 2 && elems_per_block != 5 && elems_per_block != 10)
      throw std::invalid_argument("elems_per_block must be 2, 5 or 10");
    b->scatter = scatter;
    b->elems_per_block = elems_per_block;
    b->use_graph = use_graph != 0;
    b->clear_graphs();
  }, b ? b->stream : nullptr);
}

int sfx_batch_run(sfx_batch *b) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    b->reset();
    hipEvent_t a, e;
    SFX_HIP(hipEventCreate(&a));
    SFX_HIP(hipEventCreate(&e));
    SFX_HIP(hipEventRecord(a, b->stream));
    b->advance(0, b->run_length());
    SFX_HIP(hipEventRecord(e, b->stream));
    SFX_HIP(hipEventSynchronize(e));
    float ms = 0.0f;
    SFX_HIP(hipEventElapsedTime(&ms, a, e));
    b->last_run_ms = ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(e);
  }, b ? b->stream : nullptr);
}

int sfx_batch_get_seismograms(const sfx_batch *b, int i, float *out) {
  return guarded([&] {
    if (!b || !out) throw std::invalid_argument("null argument");
    if (i < 0 || i >= b->nsim) throw std::invalid_argument("simulation index out of range");
    SFX_HIP(hipStreamSynchronize(b->stream));
    const int k = b->pos_of[i];
    const size_t n = b->seis_off[k + 1] - b->seis_off[k];
    if (n) SFX_HIP(hipMemcpy(out, b->seismo.p + b->seis_off[k], n * sizeof(float), hipMemcpyDeviceToHost));
  }, b ? b->stream : nullptr);
}

int sfx_batch_get_info(const sfx_batch *b, double *out8) {
  return guarded([&] {
    if (!b || !out8) throw std::invalid_argument("null argument");
    double ptsteps = 0.0;
    for (int k = 0; k < b->nsim; k++) ptsteps += static_cast<double>(b->nglob[k]) * b->nstep[k];
    out8[0] = b->last_run_ms;
    out8[1] = b->elems_per_block;
    out8[2] = b->use_graph ? 1.0 : 0.0;
    out8[3] = b->ncolors;
    out8[4] = static_cast<double>(b->npts);
    out8[5] = ptsteps;
    out8[6] = b->scatter;
    out8[7] = b->launches_per_step();
  });
}

}  // extern "C"
