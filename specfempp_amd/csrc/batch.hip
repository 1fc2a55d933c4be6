// Batched 2-D engine: many independent simulations (meshes of different
// sizes, sources, time steps and step counts) advanced together.
//
// The reference runs one simulation per process (time_marching<forward>::run,
// time_marching.tpp:9-124); a batch replaces N such runs.  All simulations are
// packed into one device arena as one disconnected mesh:
//   - simulations are sorted by step count (longest first), so the ones still
//     running at any step are a prefix of the batch;
//   - elements are ordered color-major (each simulation's own coloring), then
//     by simulation, so each color's running elements are a prefix of the
//     color's range and one launch per color covers every simulation;
//   - each simulation's points start on a 1024-point boundary, so a workgroup
//     of the fused update belongs to one simulation and reads its Newmark
//     constants (dt differs per simulation) once;
//   - Stacey entries, sources and receivers are concatenated in the same order.
// A step is sources, stiffness per color, Stacey, seismograms, fused update and
// a step-counter tick.  Between two simulation end-steps the launches are the
// same every step: such a segment is captured once into a HIP graph and
// replayed.  An autotuner times the launch geometries (elements per workgroup)
// and graph use on the batch itself.
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
  // tuning
  int elems_per_block = 5;
  bool use_graph = true;
  bool atomic = false;
  double last_run_ms = 0.0;
  std::vector<hipGraphExec_t> graph_cache;  // per k (nullptr until captured)

  Fields f() const {
    const size_t stride = 2 * npts;
    return Fields{fields.p, fields.p + stride, fields.p + 2 * stride, fields.p + 3 * stride, static_cast<int>(npts)};
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
    for (int k = 0; k < nsim; k++) {
      H[k] = &engine_host(engines[order[k]]);
      R[k] = &engine_run(engines[order[k]]);
      // every element must sit in exactly one color slot (the packing below
      // maps elements through the coloring)
      const HostAssembly &h = *H[k];
      if (h.color_offsets.size() < 2 || h.color_offsets.front() != 0 || h.color_offsets.back() != h.nspec ||
          static_cast<int>(h.perm.size()) != h.nspec)
        throw std::logic_error("batch: an engine's element coloring is incomplete");
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
    for (int k = 0; k < nsim; k++) ncolors = std::max(ncolors, static_cast<int>(H[k]->color_offsets.size()) - 1);
    color_off.assign(ncolors + 1, 0);
    color_cnt.assign(ncolors, std::vector<int>(nsim + 1, 0));
    std::vector<std::vector<int>> bidx(nsim);  // compute element -> batch element, per simulation
    for (int k = 0; k < nsim; k++) bidx[k].assign(H[k]->nspec, -1);
    std::vector<int> elem_sim, elem_ce;
    for (int c = 0; c < ncolors; c++) {
      color_off[c] = static_cast<int>(elem_sim.size());
      for (int k = 0; k < nsim; k++) {
        const auto &co = H[k]->color_offsets;
        if (c + 1 < static_cast<int>(co.size()))
          for (int slot = co[c]; slot < co[c + 1]; slot++) {
            bidx[k][H[k]->perm[slot]] = static_cast<int>(elem_sim.size());
            elem_sim.push_back(k);
            elem_ce.push_back(H[k]->perm[slot]);
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
    seismo.alloc(std::max<long long>(seis_off[nsim], 1));
    fields.alloc(8 * npts);
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

  // The launches of one step with k simulations running.
  void enqueue_step(int k) {
    hipStream_t s = stream;
    Fields ff = f();
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

  void reset() {
    launch_fill(stream, fields.p, 6 * npts, 0.0f);
    launch_fill(stream, seismo.p, seismo.n, 0.0f);
    const int zero = 0;
    SFX_HIP(hipMemcpyAsync(step.p, &zero, sizeof(int), hipMemcpyHostToDevice, stream));
    SFX_HIP(hipStreamSynchronize(stream));
  }

  // Steps [first, last) (first = 0 after reset()).
  void advance(int first, int last) {
    Throttle throttle(stream);
    int istep = first;
    while (istep < last) {
      const int k = running(istep);
      int end = istep;
      while (end < last && running(end) == k) end++;
      if (k == 0) break;
      if (use_graph) {
        if (!graph_cache[k]) {
          hipGraph_t g;
          SFX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
          try {
            enqueue_step(k);
          } catch (...) {
            (void)hipStreamEndCapture(stream, &g);
            throw;
          }
          SFX_HIP(hipStreamEndCapture(stream, &g));
          hipGraphExec_t ex;
          const hipError_t err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          SFX_HIP(err);
          graph_cache[k] = ex;
        }
        for (int i = istep; i < end; i++) {
          SFX_HIP(hipGraphLaunch(graph_cache[k], stream));
          throttle.tick();
        }
      } else {
        for (int i = istep; i < end; i++) {
          enqueue_step(k);
          throttle.tick();
        }
      }
      istep = end;
    }
    SFX_HIP(hipGetLastError());
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

int sfx_batch_tune(sfx_batch *b, int steps_per_candidate, int *choice3) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    if (steps_per_candidate < 1) throw std::invalid_argument("steps_per_candidate must be >= 1");
    const int nsteps = std::min(steps_per_candidate, std::max(b->nstep[b->nsim - 1] - 2, 1));
    double best = 1e300;
    int best_e = 5;
    bool best_g = true;
    for (int e : {2, 5, 10})
      for (bool g : {false, true}) {
        b->elems_per_block = e;
        b->use_graph = g;
        b->clear_graphs();
        const double ms = b->timed_steps(nsteps);
        if (ms < best) {
          best = ms;
          best_e = e;
          best_g = g;
        }
      }
    b->elems_per_block = best_e;
    b->use_graph = best_g;
    b->clear_graphs();
    b->reset();
    if (choice3) {
      choice3[0] = best_e;
      choice3[1] = best_g ? 1 : 0;
      choice3[2] = b->ncolors;
    }
  }, b ? b->stream : nullptr);
}

int sfx_batch_set_config(sfx_batch *b, int elems_per_block, int use_graph) {
  return guarded([&] {
    if (!b) throw std::invalid_argument("null batch");
    if (elems_per_block != 2 && elems_per_block != 5 && elems_per_block != 10)
      throw std::invalid_argument("elems_per_block must be 2, 5 or 10");
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
    b->advance(0, b->nstep_max);
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

int sfx_batch_get_info(const sfx_batch *b, double *out6) {
  return guarded([&] {
    if (!b || !out6) throw std::invalid_argument("null argument");
    double ptsteps = 0.0;
    for (int k = 0; k < b->nsim; k++) ptsteps += static_cast<double>(b->nglob[k]) * b->nstep[k];
    out6[0] = b->last_run_ms;
    out6[1] = b->elems_per_block;
    out6[2] = b->use_graph ? 1.0 : 0.0;
    out6[3] = b->ncolors;
    out6[4] = static_cast<double>(b->npts);
    out6[5] = ptsteps;
  });
}

}  // extern "C"
