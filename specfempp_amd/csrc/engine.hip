// Engine object and C ABI (include/specfempp_amd.h).
//
// The engine owns every device array (the reference's assembly owns its
// views, include/compute/assembly/assembly.hpp:34-119), drives the kernels of
// kernels.hip on its own stream, and converts C++ exceptions into status
// codes at the boundary.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/specfempp_amd.h"
#include "assembly.hpp"
#include "kernels.hpp"
#include "lattice.hpp"
#include "runtime.hpp"

namespace sfx {
std::string &last_error() {
  static thread_local std::string msg;
  return msg;
}
}  // namespace sfx

using namespace sfx;

namespace {

// Element-force scratch of sfx_run_noalloc: a file-scope device array, one
// copy per GPU, held until the process ends.  In host code `scratch` names a
// host-side stand-in, not the device copy: kernels receive the address
// hipGetSymbolAddress returns for the current device, kept per device in
// scratch_addr.
constexpr size_t NOALLOC_FLOATS = static_cast<size_t>(SFX_NOALLOC_MAX_ELEMENTS) * RE_ELEM;
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
float *scratch_addr[MAX_DEVICES] = {};

// Device address of `scratch` on `device`, looked up once per device with
// that device current (the previous current device is restored).
float *scratch_on(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw std::invalid_argument("device ordinal out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_addr[device]) {
    int prev = 0;
    SFX_HIP(hipGetDevice(&prev));
    if (prev != device) SFX_HIP(hipSetDevice(device));
    void *ptr = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != device) SFX_HIP(hipSetDevice(prev));
    SFX_HIP(rc);
    if (!ptr) throw HipError("hipGetSymbolAddress returned no address for the no-alloc scratch");
    scratch_addr[device] = static_cast<float *>(ptr);
  }
  return scratch_addr[device];
}

}  // namespace

struct sfx_engine {
  HostAssembly h;  // host copy (numbering and locations are returned from here)
  sfx_run_desc run{};
  bool host_only = false;  // created by sfx_engine_create_host_*: no device state
  hipStream_t stream = nullptr;
  GllConst q{};
  NewmarkConst nc{};
  int nsig = 0;
  // device arrays
  DevBuf<int> gather, bnd_gather, bnd_slot, bnd_seg, src_order, src_ispec, rec_ispec, types;
  DevBuf<float> xix, xiz, gammax, gammaz, jac, kappa, mu;
  DevBuf<float> geo;     // 2-D affine meshes: one geometry record per element (the forward stiffness reads it)
  DevBuf<float> geo3;    // 3-D database meshes: per-point derivatives and jacobian (nspec, 10, 125)
  // 2-D gather path: element forces, the per-point inverse index, Stacey
  // entries sorted by element point with their segments
  DevBuf<float> re2;
  // 2-D gather path on meshes without constant element geometry: the control
  // nodes per element and the shape derivatives at the GLL points, from which
  // the element forces recompute each point's geometry; and (2-D and 3-D
  // gather paths) (kappa, mu) per element when both are the same at all
  // points of every element
  DevBuf<float> ctrl, dsh, kmu_e;
  DevBuf<int> inv_off, inv_ent, g2_seg, g2_gather, g2_slot;
  DevBuf<float> g2_normal, g2_weight, g2_rvp, g2_rvs;
  int g2_nseg = 0;
  DevBuf<float> bnd_normal, bnd_weight, bnd_rvp, bnd_rvs;
  DevBuf<float> src_array, stf, rec_h, seismo;
  DevBuf<float> fields;  // u, v, a, minv: 4 * ncomp * nglob_pad
  DevBuf<float> re;      // 3-D gather path: element forces (nspec, 3, 128), allocated at first use
  float *re_p = nullptr; // element-force buffer of the 3-D gather kernels: re.p, or the no-alloc scratch
  int device = 0;        // HIP device the engine's memory and stream live on
  DevBuf<float> zero;    // 384 zeros read in place of absent element contributions
  DevBuf<float> bs;      // 3-D block path: block-surface force sums (nblocks, 3, 388)
  DevBuf<float> snap;    // 3-D block path: u, v at receiver element points (nrec, 125, 6)
  DevBuf<int> rec_block; // 3-D block path: 1 for blocks holding a receiver element
  // 2-D adjoint: Stacey-element record, adjoint field, kernels, adjoint sources
  DevBuf<float> record;      // (nstep, nstacey, 25, 2)
  DevBuf<float> snapshots;   // checkpoints (nsnap, 3, ncomp, nglob_pad): u, v, a after steps 0, k, 2k, ...
  int snapshot_every = 0;
  DevBuf<int> rec_index;     // per slot: Stacey element index or -1
  DevBuf<int> bnd_rec;       // per Stacey entry: Stacey element index * 25 + point
  DevBuf<float> rho;         // per point, slot order
  DevBuf<float> adj_fields;  // u, v, a of the adjoint field
  DevBuf<float> kern;        // (6, nspec, 25) slot order
  DevBuf<float> kacc;        // accumulators of the combined run (6, nspec, 25)
  DevBuf<float> racc;        // per point: sum over steps of dt a+ . u_b
  DevBuf<float> bwd_fields;  // backward wavefield of the seam calls, (3, ncomp, nglob_pad), at first use
  DevBuf<float> buf_fields;  // buffer wavefield of the seam calls (the forward state read back)
  bool kern_dirty = false;   // the accumulators changed since kern was last finished
  DevBuf<int> adj_order, adj_slot;
  DevBuf<float> adj_array, adj_stf;
  int nstacey = 0;
  bool recording = false;    // forward steps store the Stacey-element record
  bool recorded = false;     // a complete recorded forward run is in the fields
  bool adj_ready = false;
  int rec_mode = 0;          // stiffness: 0 plain, 1 record, 2 replay
  // 2-D graph capture: a device step counter read by the step-dependent
  // kernels, and one captured graph per kind of regular step
  DevBuf<int> d_step;
  hipGraphExec_t g_fwd = nullptr, g_rec = nullptr, g_adj = nullptr;
  bool use_graphs = true;
  float *rec_step = nullptr;
  bool initialized = false;
  bool a_stale = false;  // fused gather steps leave the stored acceleration behind
  int nbseg = 0;         // Stacey segments (3-D)
  bool sources_off_boundary = false;  // 3-D gather: no source element has an absorbing face
  DevBuf<float> minv1;                // 3-D gather: M^-1 of one component (nglob_pad)
  // 2-D acoustic medium: the acoustic field (chi, chi_dot, chi_dot_dot, mass
  // inverse; nglob_ac_pad each), per-point tables in slot order, the
  // interface entries, and the sources and receivers of each medium
  DevBuf<float> ac_fields, rho_inv, ac_stacey;
  DevBuf<int> ac_dirichlet, cpl_ac_self, cpl_ac_other, cpl_el_self, cpl_el_other, src_order_ac, rec_el, rec_ac;
  DevBuf<float> cpl_ac_coef, cpl_el_coef;
  int nsrc_order_el = 0;
  Lattice lat{};
  std::vector<hipEvent_t> ev;
  double timings[4] = {0, 0, 0, 0};

  Fields f() const {
    const size_t stride = static_cast<size_t>(h.ncomp) * h.nglob_pad;
    return Fields{fields.p, fields.p + stride, fields.p + 2 * stride, fields.p + 3 * stride, h.nglob_pad};
  }
  // the acoustic field of a mesh with acoustic elements
  Fields fa() const {
    const size_t n = h.nglob_ac_pad;
    return Fields{ac_fields.p, ac_fields.p + n, ac_fields.p + 2 * n, ac_fields.p + 3 * n, h.nglob_ac_pad};
  }

  ~sfx_engine() {
    invalidate_graphs();
    for (auto e : ev) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void invalidate_graphs() {
    for (hipGraphExec_t *g : {&g_fwd, &g_rec, &g_adj})
      if (*g) {
        (void)hipGraphExecDestroy(*g);
        *g = nullptr;
      }
  }
  template <class Enqueue> hipGraphExec_t capture(Enqueue &&enqueue) {
    hipGraph_t g;
    SFX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      enqueue();
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
  void set_step(int v) {
    d_step.ensure(1);
    SFX_HIP(hipMemcpyAsync(d_step.p, &v, sizeof(int), hipMemcpyHostToDevice, stream));
  }
  bool graph_2d() const { return use_graphs && h.dim == 2 && !h.coupled(); }

  // A regular 2-D forward step with every step-dependent value read from the
  // device counter; recording adds the Stacey-element record.
  void enqueue_forward_2d(bool record_on) {
    Fields ff = f();
    if (gather2d()) {
      forces_2d(ff, 0, d_step.p);
      launch_seismograms(stream, 2, h.npe, h.nrec, run.nseismo_types, types.p, 0, rec_ispec.p, gather.p, rec_h.p, ff,
                         nc.dto2, 1, seismo.p, PointGeom2D{xix.p, xiz.p, gammax.p, gammaz.p, kappa.p, mu.p}, q,
                         d_step.p, run.nstep_between_samples, -1, fs2());
      launch_assemble_update_2d(stream, h.nglob, fs2(), ff, nc, 0);
      launch_step_tick(stream, d_step.p, 1);
      return;
    }
    const long long rstride = static_cast<long long>(nstacey) * h.npe * h.ncomp;
    launch_sources(stream, 2, h.npe, h.nsrc, 0, src_order.p, src_ispec.p, gather.p, src_array.p, stf.p, ff, d_step.p);
    const bool atomic = run.scatter == SFX_SCATTER_ATOMIC;
    for (size_t c = 0; c + 1 < h.color_offsets.size(); c++) {
      const int first = h.color_offsets[c], count = h.color_offsets[c + 1] - first;
      launch_stiffness_2d(stream, first, count, atomic, gather.p, xix.p, xiz.p, gammax.p, gammaz.p, jac.p, kappa.p,
                          mu.p, q, ff, record_on ? 1 : 0, rec_index.p, record_on ? record.p : nullptr, d_step.p,
                          rstride, geo.p);
    }
    launch_stacey(stream, 2, static_cast<int>(h.bnd_gather.size()), bnd_gather.p, bnd_normal.p, bnd_weight.p,
                  bnd_rvp.p, bnd_rvs.p, ff, bnd_rec.p, record_on ? record.p : nullptr, d_step.p, rstride);
    launch_seismograms(stream, 2, h.npe, h.nrec, run.nseismo_types, types.p, 0, rec_ispec.p, gather.p, rec_h.p, ff,
                       nc.dto2, 1, seismo.p, PointGeom2D{xix.p, xiz.p, gammax.p, gammaz.p, kappa.p, mu.p}, q,
                       d_step.p, run.nstep_between_samples);
    launch_update_fused(stream, h.ncomp, ff, nc);
    launch_step_tick(stream, d_step.p, 1);
  }

  // Per-element array in device slot order (colors contiguous).
  template <class T> std::vector<T> permuted(const std::vector<T> &v, size_t width = 0) const {
    const size_t npe = width ? width : h.npe;
    std::vector<T> out(v.size());
    for (int slot = 0; slot < h.nspec; slot++)
      std::memcpy(&out[slot * npe], &v[static_cast<size_t>(h.perm[slot]) * npe], npe * sizeof(T));
    return out;
  }
  std::vector<int32_t> slots_of(const std::vector<int32_t> &ispec) const {
    std::vector<int32_t> out(ispec.size());
    for (size_t i = 0; i < ispec.size(); i++) out[i] = h.inv_perm[ispec[i]];
    return out;
  }

  bool gather_path() const { return h.dim == 3 && run.scatter == SFX_SCATTER_GATHER; }
  bool gather2d() const { return h.dim == 2 && run.scatter == SFX_SCATTER_GATHER; }
  ForceSum2D fs2() const { return ForceSum2D{inv_off.p, inv_ent.p, re2.p}; }
  // where the 2-D gather path's element forces take each point's geometry from
  int geom2d_form() const {
    if (h.dim != 2) return GEOM2D_POINT;
    if (!h.geo2d.empty()) return GEOM2D_RECORD;
    if (gather2d() && (h.ngnod2d == 4 || h.ngnod2d == 9)) return h.ngnod2d;
    return GEOM2D_POINT;
  }
  Geom2D geom2d() const {
    Geom2D gm;
    gm.xix = xix.p;
    gm.xiz = xiz.p;
    gm.gammax = gammax.p;
    gm.gammaz = gammaz.p;
    gm.jac = jac.p;
    gm.mode = geom2d_form();
    gm.record = geo.p;
    gm.ctrl = ctrl.p;
    gm.dsh = dsh.p;
    return gm;
  }
  // (kappa, mu) per element slot for the 2-D and 3-D gather paths when every
  // element holds one value of each at all of its points (bit for bit); empty
  // otherwise
  std::vector<float> element_properties() const {
    if (!gather2d() && !gather_path()) return {};
    const size_t npe = h.npe;
    std::vector<float> kmu(2 * static_cast<size_t>(h.nspec));
    auto same = [](float a, float b) { return std::memcmp(&a, &b, sizeof(float)) == 0; };
    for (int slot = 0; slot < h.nspec; slot++) {
      const size_t b = static_cast<size_t>(h.perm[slot]) * npe;
      for (size_t p = 1; p < npe; p++)
        if (!same(h.kappa[b + p], h.kappa[b]) || !same(h.mu[b + p], h.mu[b])) return {};
      kmu[2 * static_cast<size_t>(slot)] = h.kappa[b];
      kmu[2 * static_cast<size_t>(slot) + 1] = h.mu[b];
    }
    return kmu;
  }
  void upload_element_properties(hipStream_t s) {
    const std::vector<float> kmu = element_properties();
    if (kmu.empty())
      kmu_e.alloc(0);
    else
      kmu_e.upload(kmu, s);
  }

  // 2-D gather path: per field point the re2 offsets of its element
  // contributions (ascending element slot, then point), and the Stacey entries
  // sorted by element point, one segment per point
  void upload_gather_2d(hipStream_t s) {
    const int npe = h.npe;
    std::vector<int> off(static_cast<size_t>(h.nglob) + 1, 0);
    for (int slot = 0; slot < h.nspec; slot++)
      for (int p = 0; p < npe; p++) off[h.gather[static_cast<size_t>(h.perm[slot]) * npe + p] + 1]++;
    for (int g = 0; g < h.nglob; g++) off[g + 1] += off[g];
    std::vector<int> ent(off.back()), fill(off.begin(), off.end() - 1);
    for (int slot = 0; slot < h.nspec; slot++)
      for (int p = 0; p < npe; p++) {
        const int g = h.gather[static_cast<size_t>(h.perm[slot]) * npe + p];
        ent[fill[g]++] = slot * RE2_ELEM + 2 * p;
      }
    inv_off.upload(off, s);
    inv_ent.upload(ent, s);
    re2.alloc(static_cast<size_t>(h.nspec) * RE2_ELEM);
    launch_fill(s, re2.p, re2.n, 0.0f);
    const size_t nb = h.bnd_gather.size();
    std::vector<size_t> order(nb);
    std::iota(order.begin(), order.end(), 0);
    auto dev_slot = [&](size_t q) {
      const int c = h.bnd_slot[q] / npe, p = h.bnd_slot[q] - (h.bnd_slot[q] / npe) * npe;
      return h.inv_perm[c] * npe + p;
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return dev_slot(a) < dev_slot(b); });
    std::vector<int> sg, gg(nb), sl(nb);
    std::vector<float> nrm(2 * nb), wt(nb), rvp(nb), rvs(nb);
    for (size_t k = 0; k < nb; k++) {
      const size_t q = order[k];
      gg[k] = h.bnd_gather[q];
      sl[k] = dev_slot(q);
      nrm[2 * k] = h.bnd_normal[2 * q];
      nrm[2 * k + 1] = h.bnd_normal[2 * q + 1];
      wt[k] = h.bnd_weight[q];
      rvp[k] = h.bnd_rho_vp[q];
      rvs[k] = h.bnd_rho_vs[q];
      if (k == 0 || sl[k] != sl[k - 1]) sg.push_back(static_cast<int>(k));
    }
    sg.push_back(static_cast<int>(nb));
    g2_nseg = static_cast<int>(sg.size()) - 1;
    g2_seg.upload(sg, s);
    g2_gather.upload(gg, s);
    g2_slot.upload(sl, s);
    g2_normal.upload(nrm, s);
    g2_weight.upload(wt, s);
    g2_rvp.upload(rvp, s);
    g2_rvs.upload(rvs, s);
  }
  // element forces, Stacey and sources into re2 (the step's force, unassembled)
  void forces_2d(Fields ff, int istep, const int *step) {
    launch_elemforce_2d(stream, h.nspec, gather.p, geom2d(), kappa.p, mu.p, kmu_e.p, q, ff.u, h.nglob_pad, re2.p);
    launch_stacey_re2(stream, g2_nseg, g2_seg.p, g2_gather.p, g2_slot.p, g2_normal.p, g2_weight.p, g2_rvp.p, g2_rvs.p,
                      ff, re2.p);
    launch_sources_re2(stream, h.nsrc, istep, src_order.p, src_ispec.p, src_array.p, stf.p, re2.p, step);
  }
  bool block_path() const { return h.dim == 3 && run.scatter == SFX_SCATTER_BLOCK; }

  // Stacey entries regrouped for the block path: entry slot = owner * 386 +
  // slot of its point for the block of its element (lattice.hpp), entries
  // sorted (stably) by that slot, one segment per slot.
  void upload_block_boundary(hipStream_t s) {
    const size_t nb = h.bnd_gather.size();
    const int nbx = h.nx / 2, nby = h.ny / 2;
    std::vector<int32_t> bslot(nb);
    for (size_t q = 0; q < nb; q++) {
      const int e = h.bnd_slot[q] / 125, p = h.bnd_slot[q] % 125;
      const int ex = e % h.nx, ey = (e / h.nx) % h.ny, ez = e / (h.nx * h.ny);
      const int bi = 4 * (ex & 1) + p % 5, bj = 4 * (ey & 1) + (p / 5) % 5, bk = 4 * (ez & 1) + p / 25;
      int owner, slot;
      if (!block_point_owner(nbx, nby, ex / 2, ey / 2, ez / 2, bi, bj, bk, owner, slot))
        throw std::logic_error("internal: Stacey point inside a block");
      bslot[q] = owner * OWNER_SLOTS + slot;
    }
    std::vector<int32_t> order(nb);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bslot[a] < bslot[b]; });
    std::vector<int32_t> g(nb), sl(nb), seg;
    std::vector<float> nrm(3 * nb), wt(nb), rvp(nb), rvs(nb);
    for (size_t q = 0; q < nb; q++) {
      const int o = order[q];
      g[q] = h.bnd_gather[o];
      sl[q] = bslot[o];
      for (int d = 0; d < 3; d++) nrm[3 * q + d] = h.bnd_normal[3 * o + d];
      wt[q] = h.bnd_weight[o];
      rvp[q] = h.bnd_rho_vp[o];
      rvs[q] = h.bnd_rho_vs[o];
      if (q == 0 || sl[q] != sl[q - 1]) seg.push_back(static_cast<int32_t>(q));
    }
    seg.push_back(static_cast<int32_t>(nb));
    bnd_gather.upload(g, s);
    bnd_slot.upload(sl, s);
    bnd_seg.upload(seg, s);
    bnd_normal.upload(nrm, s);
    bnd_weight.upload(wt, s);
    bnd_rvp.upload(rvp, s);
    bnd_rvs.upload(rvs, s);
    nbseg = static_cast<int>(seg.size()) - 1;
    std::vector<int32_t> rb(static_cast<size_t>(nbx) * nby * (h.nz / 2), 0);
    for (int e : h.rec_ispec) {
      const int ex = e % h.nx, ey = (e / h.nx) % h.ny, ez = e / (h.nx * h.ny);
      rb[(ez / 2 * nby + ey / 2) * nbx + ex / 2] = 1;
    }
    rec_block.upload(rb, s);
    bs.alloc(block3d_surface_floats(lat));
    launch_fill(s, bs.p, bs.n, 0.0f);
    snap.alloc(static_cast<size_t>(h.nrec) * 125 * 6);
  }

  void upload() {
    lat = Lattice{h.nx, h.ny, h.nz, h.nx + 1, h.ny + 1, h.nz + 1, h.nint_pad};
    SFX_HIP(hipGetDevice(&device));
    SFX_HIP(hipStreamCreate(&stream));
    hipStream_t s = stream;
    gather.upload(permuted(h.gather), s);
    kappa.upload(permuted(h.kappa), s);
    mu.upload(permuted(h.mu), s);
    upload_element_properties(s);
    if (h.general3d) geo3.upload(permuted(h.geo3, 10 * static_cast<size_t>(h.npe)), s);
    if (h.dim == 2) {
      xix.upload(permuted(h.xix), s);
      xiz.upload(permuted(h.xiz), s);
      gammax.upload(permuted(h.gammax), s);
      gammaz.upload(permuted(h.gammaz), s);
      jac.upload(permuted(h.jac), s);
      if (!h.geo2d.empty()) geo.upload(permuted(h.geo2d, 8), s);
      if (geom2d_form() == GEOM2D_CTRL4 || geom2d_form() == GEOM2D_CTRL9) {
        ctrl.upload(permuted(h.ctrl2d, 2 * static_cast<size_t>(h.ngnod2d)), s);
        dsh.upload(h.dsh2d, s);
      }
      rho.upload(permuted(h.rho), s);
      // Stacey elements are the last ones in compute order
      nstacey = h.nspec - h.nspec_none;
      std::vector<int32_t> ri(h.nspec), br(h.bnd_slot.size());
      for (int slot = 0; slot < h.nspec; slot++) ri[slot] = h.perm[slot] >= h.nspec_none ? h.perm[slot] - h.nspec_none : -1;
      for (size_t q = 0; q < br.size(); q++) {
        const int c = h.bnd_slot[q] / h.npe, lp = h.bnd_slot[q] % h.npe;
        if (c < h.nspec_none) throw std::logic_error("internal: Stacey entry outside the Stacey elements");
        br[q] = (c - h.nspec_none) * h.npe + lp;
      }
      rec_index.upload(ri, s);
      bnd_rec.upload(br, s);
      if (gather2d()) upload_gather_2d(s);
    }
    bnd_gather.upload(h.bnd_gather, s);
    bnd_slot.upload(h.bnd_slot, s);
    bnd_seg.upload(h.bnd_seg, s);
    nbseg = static_cast<int>(h.bnd_seg.size()) - 1;
    if (gather_path()) {
      zero.alloc(384);
      launch_fill(s, zero.p, zero.n, 0.0f);
      // M^-1 of one component for the tiles without absorbing-face points,
      // where the three components are equal
      {
        const size_t np = h.nglob_pad;
        std::vector<float> m1(h.mass_inverse.begin(), h.mass_inverse.begin() + np);
        for (int tz = 0; tz < lat.TZ; tz++)
          for (int ty = 0; ty < lat.TY; ty++)
            for (int tx = 0; tx < lat.TX; tx++) {
              const int am = h.absorb_mask;
              if (((am & 1) && tx == 0) || ((am & 2) && tx == h.nx) || ((am & 4) && ty == 0) ||
                  ((am & 8) && ty == h.ny) || ((am & 16) && tz == 0) || ((am & 32) && tz == h.nz))
                continue;
              const size_t t0 = ((static_cast<size_t>(tz) * lat.TY + ty) * lat.TX + tx) * 64;
              for (size_t q = t0; q < t0 + 64; q++)
                if (h.mass_inverse[np + q] != m1[q] || h.mass_inverse[2 * np + q] != m1[q])
                  throw std::logic_error("internal: M^-1 components differ off the absorbing faces");
            }
        minv1.upload(m1, s);
      }
      // sources and Stacey share one launch when no source element has an
      // absorbing face (their entries are then disjoint)
      std::vector<char> absorbing(h.nspec, 0);
      for (int32_t sl : h.bnd_slot) absorbing[sl / 125] = 1;
      sources_off_boundary = true;
      for (int32_t sl : slots_of(h.src_ispec))
        if (sl < 0 || sl >= h.nspec || absorbing[sl]) sources_off_boundary = false;
    }
    bnd_normal.upload(h.bnd_normal, s);
    bnd_weight.upload(h.bnd_weight, s);
    bnd_rvp.upload(h.bnd_rho_vp, s);
    bnd_rvs.upload(h.bnd_rho_vs, s);
    if (block_path()) upload_block_boundary(s);
    if (h.coupled()) {
      src_order.upload(h.src_order_el, s);
      nsrc_order_el = static_cast<int>(h.src_order_el.size());
      src_order_ac.upload(h.src_order_ac, s);
      rec_el.upload(h.rec_el, s);
      rec_ac.upload(h.rec_ac, s);
      rho_inv.upload(permuted(h.rho_inv), s);
      ac_stacey.upload(permuted(h.ac_stacey), s);
      ac_dirichlet.upload(permuted(h.ac_dirichlet), s);
      cpl_ac_self.upload(h.cpl_ac_self, s);
      cpl_ac_other.upload(h.cpl_ac_other, s);
      cpl_ac_coef.upload(h.cpl_ac_coef, s);
      cpl_el_self.upload(h.cpl_el_self, s);
      cpl_el_other.upload(h.cpl_el_other, s);
      cpl_el_coef.upload(h.cpl_el_coef, s);
      const size_t na = h.nglob_ac_pad;
      if (h.mass_inverse_ac.size() != na) throw std::logic_error("internal: acoustic mass inverse layout");
      ac_fields.alloc(4 * na);
      launch_fill(s, ac_fields.p, 3 * na, 0.0f);
      SFX_HIP(hipMemcpyAsync(ac_fields.p + 3 * na, h.mass_inverse_ac.data(), na * sizeof(float),
                             hipMemcpyHostToDevice, s));
    } else {
      src_order.upload(h.src_order, s);
      nsrc_order_el = h.nsrc;
    }
    src_ispec.upload(slots_of(h.src_ispec), s);
    src_array.upload(h.src_array, s);
    stf.upload(h.stf, s);
    rec_ispec.upload(slots_of(h.rec_ispec), s);
    rec_h.upload(h.rec_h, s);
    std::vector<int> t(run.seismo_types, run.seismo_types + run.nseismo_types);
    types.upload(t, s);
    seismo.alloc(static_cast<size_t>(nsig) * run.nseismo_types * h.nrec * h.ncomp);
    if (seismo.n) launch_fill(s, seismo.p, seismo.n, 0.0f);
    const size_t stride = static_cast<size_t>(h.ncomp) * h.nglob_pad;
    fields.alloc(4 * stride);
    launch_fill(s, fields.p, 4 * stride, 0.0f);
    // mass inverse, already in device layout (padding slots hold 0)
    if (h.mass_inverse.size() != stride) throw std::logic_error("internal: mass inverse layout");
    SFX_HIP(hipMemcpyAsync(fields.p + 3 * stride, h.mass_inverse.data(), stride * sizeof(float),
                           hipMemcpyHostToDevice, s));
    SFX_HIP(hipStreamSynchronize(s));
    // quadrature constants
    for (int i = 0; i < 5; i++) {
      q.w[i] = h.w[i];
      for (int l = 0; l < 5; l++) {
        q.hp[i][l] = h.hprime[i][l];
        q.hpw[i][l] = h.hprime[l][i] * h.w[l];
      }
    }
    const float dt = static_cast<float>(run.dt);
    nc.dt = dt;
    nc.dto2 = dt / 2.0;
    nc.dt2o2 = dt * dt / 2.0;
  }

  // The 3-D gather kernels of the allocating entry points use the engine's
  // own element-force buffer, allocated here at first use (3 components x 128
  // floats per element, 125 used).
  void own_forces() {
    if (!gather_path()) return;
    if (!re.p) {
      re.alloc(static_cast<size_t>(h.nspec) * RE_ELEM);
      launch_fill(stream, re.p, re.n, 0.0f);
    }
    re_p = re.p;
  }
  // Gather path: element forces into re, then sources and Stacey into re.
  void element_forces(Fields ff, int istep) {
    if (istep < 0 || istep >= run.nstep) throw std::invalid_argument("istep out of range");
    launch_elemforce_3d(stream, lat, kappa.p, mu.p, kmu_e.p, h.xix3, h.ety3, h.gaz3, h.jac3, q, ff.u, h.nglob_pad,
                        re_p);
  }
  void element_boundary_sources(Fields ff, int istep) {
    if (sources_off_boundary) {
      launch_stacey_sources_re(stream, nbseg, bnd_seg.p, bnd_gather.p, bnd_slot.p, bnd_normal.p, bnd_weight.p,
                               bnd_rvp.p, bnd_rvs.p, ff, h.nsrc, istep, src_ispec.p, src_array.p, stf.p, re_p);
      return;
    }
    launch_sources_re(stream, h.nsrc, istep, src_ispec.p, src_array.p, stf.p, re_p);
    launch_stacey_re(stream, nbseg, bnd_seg.p, bnd_gather.p, bnd_slot.p, bnd_normal.p, bnd_weight.p, bnd_rvp.p,
                     bnd_rvs.p, ff, re_p);
  }
  // Block path: element forces, sources and block-interior assembly, then
  // Stacey into the block-surface sums.
  void block_forces(Fields ff, int istep, int store_a, int divide_only) {
    if (istep < 0 || istep >= run.nstep) throw std::invalid_argument("istep out of range");
    launch_block_step_3d(stream, lat, kappa.p, mu.p, h.xix3, h.ety3, h.gaz3, h.jac3, q, ff, nc, h.nsrc, istep,
                         src_ispec.p, src_array.p, stf.p, rec_block.p, store_a, divide_only, bs.p);
  }
  void block_boundary(Fields ff) {
    launch_stacey_bs(stream, nbseg, bnd_seg.p, bnd_gather.p, bnd_slot.p, bnd_normal.p, bnd_weight.p, bnd_rvp.p,
                     bnd_rvs.p, ff, bs.p);
  }
  // Bring the stored acceleration up to date after fused gather steps.
  void refresh_acceleration() {
    if (gather_path() && a_stale) {
      launch_assemble_3d(stream, lat, re_p, f(), 1);
      a_stale = false;
    }
    if (gather2d() && a_stale) {  // re2 still holds the last step's forces
      launch_assemble_update_2d(stream, h.nglob, fs2(), f(), nc, 1);
      a_stale = false;
    }
  }

  // One launch per color (plain read-modify-write), or one atomic launch.
  void stiffness(Fields ff) {
    const bool atomic = run.scatter == SFX_SCATTER_ATOMIC;
    for (size_t c = 0; c < static_cast<size_t>(h.ncolors_el); c++) {
      const int first = h.color_offsets[c], count = h.color_offsets[c + 1] - first;
      if (h.dim == 2)
        launch_stiffness_2d(stream, first, count, atomic, gather.p, xix.p, xiz.p, gammax.p, gammaz.p, jac.p, kappa.p,
                            mu.p, q, ff, rec_mode, rec_index.p, rec_step, nullptr, 0, geo.p);
      else if (h.general3d)
        launch_stiffness_3d_general(stream, first, count, atomic, gather.p, geo3.p, kappa.p, mu.p, q, ff);
      else
        launch_stiffness_3d(stream, first, count, atomic, gather.p, kappa.p, mu.p, h.xix3, h.ety3, h.gaz3, h.jac3, q,
                            ff);
    }
  }
  void boundary(Fields ff) {
    launch_stacey(stream, h.dim, static_cast<int>(h.bnd_gather.size()), bnd_gather.p, bnd_normal.p, bnd_weight.p,
                  bnd_rvp.p, bnd_rvs.p, ff, bnd_rec.p, rec_mode == 1 ? rec_step : nullptr);
  }
  // the sources of the elastic medium (all sources without acoustic elements)
  void sources(Fields ff, int istep) {
    if (istep < 0 || istep >= run.nstep) throw std::invalid_argument("istep out of range");
    launch_sources(stream, h.dim, h.npe, h.nsrc, istep, src_order.p, src_ispec.p, gather.p, src_array.p, stf.p, ff,
                   nullptr, nsrc_order_el);
  }
  void seismograms(Fields ff, int isig, int corrected) { seismograms(ff, isig, corrected, nc.dto2); }
  // Seismograms of the elastic receivers; with acoustic elements also those
  // of the acoustic receivers, from the acoustic field's current state.
  void seismograms(Fields ff, int isig, int corrected, float dto2) {
    if (isig < 0 || isig >= nsig) throw std::invalid_argument("seismogram step out of range");
    const bool cpl = h.coupled();
    launch_seismograms(stream, h.dim, h.npe, h.nrec, run.nseismo_types, types.p, isig, rec_ispec.p, gather.p, rec_h.p,
                       ff, dto2, corrected, seismo.p, PointGeom2D{xix.p, xiz.p, gammax.p, gammaz.p, kappa.p, mu.p},
                       q, nullptr, 1, -1, ForceSum2D{}, cpl ? rec_el.p : nullptr,
                       cpl ? static_cast<int>(h.rec_el.size()) : -1);
    if (cpl)
      launch_acoustic_seismograms(stream, static_cast<int>(h.rec_ac.size()), rec_ac.p, h.nrec, run.nseismo_types,
                                  types.p, isig, rec_ispec.p, gather.p, rec_h.p, xix.p, xiz.p, gammax.p, gammaz.p,
                                  rho_inv.p, q, fa(), seismo.p);
  }
  // ---- acoustic medium and coupling (meshes with acoustic elements) ----
  // acoustic colors: color_offsets[ncolors_el ..]
  void acoustic_stiffness(Fields fac) {
    const bool atomic = run.scatter == SFX_SCATTER_ATOMIC;
    for (size_t c = h.ncolors_el; c + 1 < h.color_offsets.size(); c++) {
      const int first = h.color_offsets[c], count = h.color_offsets[c + 1] - first;
      launch_acoustic_stiffness_2d(stream, first, count, atomic, gather.p, xix.p, xiz.p, gammax.p, gammaz.p, jac.p,
                                   rho_inv.p, ac_stacey.p, ac_dirichlet.p, q, fac);
    }
  }
  // update_wavefields<acoustic> (domain_kernels.hpp:69-126): coupling from the
  // elastic displacement, sources, stiffness with Stacey and free surface;
  // then the mass division
  void update_acoustic(int istep, int correct = 0) {
    if (istep < 0 || istep >= run.nstep) throw std::invalid_argument("istep out of range");
    Fields fac = fa();
    launch_coupling_2d(stream, static_cast<int>(h.cpl_ac_self.size()), cpl_ac_self.p, cpl_ac_other.p, cpl_ac_coef.p,
                       fac, f(), 0);
    launch_acoustic_sources(stream, static_cast<int>(h.src_order_ac.size()), h.nsrc, istep, src_order_ac.p,
                            src_ispec.p, gather.p, src_array.p, stf.p, kappa.p, fac);
    acoustic_stiffness(fac);
    launch_divide_corrector(stream, 1, fac, nc, 1, correct);
  }
  // the elastic side's interface term, before the elastic sources
  void couple_into_elastic(Fields ff) {
    launch_coupling_2d(stream, static_cast<int>(h.cpl_el_self.size()), cpl_el_self.p, cpl_el_other.p, cpl_el_coef.p,
                       ff, fa(), 1);
  }
  void require_init() const {
    if (!initialized) throw std::logic_error("sfx_initialize must be called first");
  }

  // One fused Newmark step; `evs` (5 events) brackets the kernel families.
  // `last`: the final step of a call; the block path then stores the
  // acceleration of every point so that the fields read back are complete.
  void step(int istep, hipEvent_t *evs, bool last) {
    Fields ff = f();
    if (block_path()) {
      const bool sample = istep % run.nstep_between_samples == 0;
      const int isig = istep / run.nstep_between_samples;
      if (sample && isig >= nsig) throw std::invalid_argument("seismogram step out of range");
      if (evs) SFX_HIP(hipEventRecord(evs[0], stream));
      if (sample) launch_receiver_snapshot(stream, lat, h.nrec, rec_ispec.p, ff, snap.p);
      if (evs) SFX_HIP(hipEventRecord(evs[1], stream));
      block_forces(ff, istep, last ? 2 : (sample ? 1 : 0), 0);
      if (evs) SFX_HIP(hipEventRecord(evs[2], stream));
      block_boundary(ff);
      if (sample)
        launch_seismograms_block(stream, lat, h.nrec, run.nseismo_types, types.p, isig, rec_ispec.p, rec_h.p, bs.p,
                                 snap.p, ff, nc.dto2, seismo.p);
      if (evs) SFX_HIP(hipEventRecord(evs[3], stream));
      launch_surface_update_3d(stream, lat, bs.p, ff, nc, last ? 1 : 0, 0);
      if (evs) SFX_HIP(hipEventRecord(evs[4], stream));
      return;
    }
    if (gather2d()) {
      if (istep < 0 || istep >= run.nstep) throw std::invalid_argument("istep out of range");
      if (recording) throw std::invalid_argument("adjoint runs need SFX_SCATTER_COLORED or _ATOMIC");
      if (evs) SFX_HIP(hipEventRecord(evs[0], stream));
      if (evs) SFX_HIP(hipEventRecord(evs[1], stream));
      forces_2d(ff, istep, nullptr);
      if (evs) SFX_HIP(hipEventRecord(evs[2], stream));
      if (istep % run.nstep_between_samples == 0) {
        const int isig = istep / run.nstep_between_samples;
        if (isig >= nsig) throw std::invalid_argument("seismogram step out of range");
        launch_seismograms(stream, 2, h.npe, h.nrec, run.nseismo_types, types.p, isig, rec_ispec.p, gather.p,
                           rec_h.p, ff, nc.dto2, 1, seismo.p,
                           PointGeom2D{xix.p, xiz.p, gammax.p, gammaz.p, kappa.p, mu.p}, q, nullptr, 1, -1, fs2());
      }
      if (evs) SFX_HIP(hipEventRecord(evs[3], stream));
      launch_assemble_update_2d(stream, h.nglob, fs2(), ff, nc, 0);
      if (evs) SFX_HIP(hipEventRecord(evs[4], stream));
      a_stale = true;
      return;
    }
    if (gather_path()) {
      if (evs) SFX_HIP(hipEventRecord(evs[0], stream));
      if (evs) SFX_HIP(hipEventRecord(evs[1], stream));
      element_forces(ff, istep);
      if (evs) SFX_HIP(hipEventRecord(evs[2], stream));
      const bool sample = istep % run.nstep_between_samples == 0;
      const int isig = istep / run.nstep_between_samples;
      if (sample && isig >= nsig) throw std::invalid_argument("seismogram step out of range");
      bool need_a = false;
      for (int it = 0; it < run.nseismo_types; it++) need_a = need_a || run.seismo_types[it] != SFX_DISPLACEMENT;
      if (sample && !need_a && sources_off_boundary) {
        // Stacey, sources and displacement samples in one launch
        launch_stacey_sources_seismo_re(stream, nbseg, bnd_seg.p, bnd_gather.p, bnd_slot.p, bnd_normal.p,
                                        bnd_weight.p, bnd_rvp.p, bnd_rvs.p, ff, h.nsrc, istep, src_ispec.p,
                                        src_array.p, stf.p, re_p, lat, h.nrec, run.nseismo_types, types.p, isig,
                                        rec_ispec.p, rec_h.p, nc.dto2, seismo.p);
      } else {
        element_boundary_sources(ff, istep);
        if (sample)
          launch_seismograms_re(stream, lat, h.nrec, run.nseismo_types, types.p, isig, rec_ispec.p, rec_h.p, re_p, ff,
                                nc.dto2, seismo.p, need_a);
      }
      if (evs) SFX_HIP(hipEventRecord(evs[3], stream));
      launch_assemble_update_3d(stream, lat, re_p, zero.p, ff, nc, minv1.p, h.absorb_mask);
      if (evs) SFX_HIP(hipEventRecord(evs[4], stream));
      a_stale = true;
      return;
    }
    if (h.coupled()) {
      // time_marching.tpp:33-84 on a mesh with both media: the acoustic
      // update and corrector, then the elastic update with the interface term
      // from the corrected chi_dot_dot, seismograms of both media, and the
      // next step's predictors (the elastic one fused with its corrector)
      if (recording) throw std::invalid_argument("adjoint runs are implemented for elastic meshes");
      if (evs) SFX_HIP(hipEventRecord(evs[0], stream));
      update_acoustic(istep, 1);
      couple_into_elastic(ff);
      sources(ff, istep);
      if (evs) SFX_HIP(hipEventRecord(evs[1], stream));
      stiffness(ff);
      if (evs) SFX_HIP(hipEventRecord(evs[2], stream));
      boundary(ff);
      if (istep % run.nstep_between_samples == 0) seismograms(ff, istep / run.nstep_between_samples, 1);
      if (evs) SFX_HIP(hipEventRecord(evs[3], stream));
      launch_update_fused(stream, h.ncomp, ff, nc);
      launch_predictor(stream, 1, fa(), nc);
      if (evs) SFX_HIP(hipEventRecord(evs[4], stream));
      return;
    }
    if (recording) {
      rec_mode = 1;
      rec_step = record.p + static_cast<size_t>(istep) * nstacey * h.npe * h.ncomp;
    }
    if (evs) SFX_HIP(hipEventRecord(evs[0], stream));
    sources(ff, istep);
    if (evs) SFX_HIP(hipEventRecord(evs[1], stream));
    stiffness(ff);
    if (evs) SFX_HIP(hipEventRecord(evs[2], stream));
    boundary(ff);
    if (istep % run.nstep_between_samples == 0) seismograms(ff, istep / run.nstep_between_samples, 1);
    if (evs) SFX_HIP(hipEventRecord(evs[3], stream));
    // a recorded run ends on the corrected state of its last step, the state
    // the backward field starts from; checkpoint steps keep their corrected
    // state too
    const bool snap = recording && is_checkpoint(istep);
    if (recording && (last || snap)) {
      launch_divide_corrector(stream, h.ncomp, ff, nc, 1, 1);
      if (snap) {
        const size_t n3 = 3 * static_cast<size_t>(h.ncomp) * h.nglob_pad;
        SFX_HIP(hipMemcpyAsync(snapshots.p + static_cast<size_t>(istep / snapshot_every) * n3, fields.p,
                               n3 * sizeof(float), hipMemcpyDeviceToDevice, stream));
      }
      if (!last) launch_predictor(stream, h.ncomp, ff, nc);
    } else {
      launch_update_fused(stream, h.ncomp, ff, nc);
    }
    if (evs) SFX_HIP(hipEventRecord(evs[4], stream));
    rec_mode = 0;
    rec_step = nullptr;
  }

  // ---- wavefield / medium / direction seams (time_scheme, domain_kernels<W>) ----
  // Fields of a simulation wavefield (specfem::wavefield::simulation_field).
  // The adjoint, backward and buffer fields are allocated and zeroed at first
  // use; all share the forward field's mass matrix inverse.
  Fields wavefield(int w) {
    if (w == SFX_WAVEFIELD_FORWARD) return f();
    if (h.dim != 2) throw std::invalid_argument("3-D engines have only the forward wavefield");
    DevBuf<float> *b = w == SFX_WAVEFIELD_ADJOINT    ? &adj_fields
                       : w == SFX_WAVEFIELD_BACKWARD ? &bwd_fields
                       : w == SFX_WAVEFIELD_BUFFER   ? &buf_fields
                                                     : nullptr;
    if (!b) throw std::invalid_argument("wavefield must be one of SFX_WAVEFIELD_*");
    const size_t stride = static_cast<size_t>(h.ncomp) * h.nglob_pad;
    if (b->n != 3 * stride) {
      b->alloc(3 * stride);
      launch_fill(stream, b->p, b->n, 0.0f);
    }
    return Fields{b->p, b->p + stride, b->p + 2 * stride, fields.p + 3 * stride, h.nglob_pad};
  }
  // Elements and degrees of freedom of a medium (this engine's meshes are
  // elastic P-SV / 3-D elastic: other media hold none, and their phases
  // update nothing, as the reference's do for media absent from a mesh).
  static void check_medium(int medium) {
    if (medium < SFX_MEDIUM_ELASTIC_PSV || medium > SFX_MEDIUM_ELASTIC_PSV_T)
      throw std::invalid_argument("medium must be one of SFX_MEDIUM_*");
  }
  int medium_elements(int medium) const {
    if (medium == SFX_MEDIUM_ELASTIC_PSV) return h.nspec_el;
    if (medium == SFX_MEDIUM_ACOUSTIC) return h.nspec - h.nspec_el;
    return 0;
  }
  int medium_dofs(int medium) const {
    if (medium == SFX_MEDIUM_ELASTIC_PSV) return h.nglob_el * h.ncomp;
    if (medium == SFX_MEDIUM_ACOUSTIC) return h.nglob_ac;
    return 0;
  }
  // the fields of (medium, wavefield): acoustic fields exist for the forward
  // wavefield of meshes with acoustic elements
  Fields medium_wavefield(int medium, int w, int &ncomp) {
    if (medium == SFX_MEDIUM_ACOUSTIC) {
      if (w != SFX_WAVEFIELD_FORWARD) throw std::invalid_argument("acoustic media have only the forward wavefield");
      ncomp = 1;
      return fa();
    }
    ncomp = h.ncomp;
    return wavefield(w);
  }
  NewmarkConst newmark(int direction) const {
    if (direction == SFX_DIRECTION_FORWARD) return nc;
    if (direction == SFX_DIRECTION_BACKWARD) return NewmarkConst{-nc.dt, -nc.dto2, nc.dt2o2};  // newmark.tpp:204-242
    throw std::invalid_argument("direction must be SFX_DIRECTION_FORWARD or _BACKWARD");
  }
  void require_adjoint_capable() const {
    if (h.dim != 2) throw std::invalid_argument("adjoint and backward wavefields are implemented for 2-D meshes");
    if (h.coupled()) throw std::invalid_argument("adjoint runs are implemented for elastic meshes");
    if (gather2d()) throw std::invalid_argument("adjoint runs need SFX_SCATTER_COLORED or _ATOMIC");
  }
  // update_wavefields of the forward field (domain_kernels.hpp:69-126):
  // sources, stiffness + Stacey, mass division.
  void update_forward(int istep) {
    Fields ff = f();
    if (block_path()) {
      block_forces(ff, istep, 0, 1);
      block_boundary(ff);
      launch_surface_update_3d(stream, lat, bs.p, ff, nc, 0, 1);
    } else if (gather_path()) {
      own_forces();
      element_forces(ff, istep);
      element_boundary_sources(ff, istep);
      launch_assemble_3d(stream, lat, re_p, ff, 1);
      a_stale = false;
    } else if (gather2d()) {
      if (istep < 0 || istep >= run.nstep) throw std::invalid_argument("istep out of range");
      forces_2d(ff, istep, nullptr);
      launch_assemble_update_2d(stream, h.nglob, fs2(), ff, nc, 1);
      a_stale = false;
    } else {
      if (h.coupled()) couple_into_elastic(ff);
      sources(ff, istep);
      stiffness(ff);
      boundary(ff);
      launch_divide_corrector(stream, h.ncomp, ff, nc, 1, 0);
    }
  }
  // update_wavefields of the adjoint field: adjoint sources, stiffness and
  // Stacey traction as in a forward step.
  void update_adjoint(int istep) {
    require_adjoint_capable();
    if (istep < 0 || istep >= run.nstep) throw std::invalid_argument("istep out of range");
    Fields adj = wavefield(SFX_WAVEFIELD_ADJOINT);
    if (adj_ready)
      launch_sources(stream, 2, h.npe, 2 * h.nadj, istep, adj_order.p, adj_slot.p, gather.p, adj_array.p, adj_stf.p,
                     adj);
    stiffness(adj);
    boundary(adj);
    launch_divide_corrector(stream, h.ncomp, adj, nc, 1, 0);
  }
  // update_wavefields of the backward field: forward sources, stiffness with
  // the Stacey elements replaying the recorded forward forces of this step
  // (compute_stiffness_interaction.tpp:119-136), no traction, mass division.
  void update_backward(int istep) {
    require_adjoint_capable();
    if (istep < 0 || istep >= run.nstep) throw std::invalid_argument("istep out of range");
    const size_t rec_stride = static_cast<size_t>(nstacey) * h.npe * h.ncomp;
    if (nstacey > 0 && record.n != static_cast<size_t>(run.nstep) * rec_stride)
      throw std::logic_error("no Stacey record: run sfx_forward_record before stepping the backward wavefield");
    Fields bwd = wavefield(SFX_WAVEFIELD_BACKWARD);
    sources(bwd, istep);
    rec_mode = 2;
    rec_step = nstacey > 0 ? record.p + static_cast<size_t>(istep) * rec_stride : nullptr;
    try {
      stiffness(bwd);
    } catch (...) {
      rec_mode = 0;
      rec_step = nullptr;
      throw;
    }
    rec_mode = 0;
    rec_step = nullptr;
    launch_divide_corrector(stream, h.ncomp, bwd, nc, 1, 0);
  }
  void frechet_zero() {
    const size_t npts = static_cast<size_t>(h.nspec) * h.npe;
    kacc.ensure(6 * npts);
    launch_fill(stream, kacc.p, kacc.n, 0.0f);
    racc.ensure(h.nglob_pad);
    launch_fill(stream, racc.p, racc.n, 0.0f);
    kern_dirty = true;
  }
  void frechet_finish() {
    const size_t npts = static_cast<size_t>(h.nspec) * h.npe;
    kern.ensure(6 * npts);
    launch_frechet_finish(stream, static_cast<int>(npts), gather.p, kappa.p, mu.p, rho.p, racc.p, kacc.p + npts,
                          kacc.p + 2 * npts, kern.p);
    kern_dirty = false;
  }

  bool is_checkpoint(int istep) const {
    return snapshot_every > 0 && istep % snapshot_every == 0 && istep < run.nstep - 1;
  }

  // Combined adjoint + backward loop (time_marching.tpp:126-247), 2-D.  The
  // backward field is the forward field itself: it holds the forward run's
  // final state, which is where the reference's backward field stands after
  // its first step (the state read back is copied in at istep = nstep - 1), so
  // that step is not run.  Between steps both fields hold their predicted
  // state.  A step: sources, both fields' forces in one launch per color (with
  // the strain parts of the Fréchet kernels), the adjoint's Stacey traction,
  // backward seismograms, and one fused pass of division, correctors, the rho
  // kernel's product and the next predictors.  At a checkpoint step the
  // backward state is reset to the forward snapshot after its corrector, so
  // that step runs the phases separately.
  void adjoint_run() {
    const size_t stride = static_cast<size_t>(h.ncomp) * h.nglob_pad;
    adj_fields.ensure(3 * stride);
    launch_fill(stream, adj_fields.p, adj_fields.n, 0.0f);
    const size_t npts = static_cast<size_t>(h.nspec) * h.npe;
    kern.ensure(6 * npts);
    kacc.ensure(6 * npts);
    launch_fill(stream, kacc.p, kacc.n, 0.0f);
    racc.ensure(h.nglob_pad);
    launch_fill(stream, racc.p, racc.n, 0.0f);
    if (seismo.n) launch_fill(stream, seismo.p, seismo.n, 0.0f);
    Fields bwd = f();
    Fields adj{adj_fields.p, adj_fields.p + stride, adj_fields.p + 2 * stride, bwd.minv, h.nglob_pad};
    float *Kmu = kacc.p + npts, *Kkappa = kacc.p + 2 * npts;
    const float dt = nc.dt;
    const NewmarkConst back{-dt, -nc.dto2, nc.dt2o2};  // newmark.tpp:204-242
    const int nstep = run.nstep, nsb = run.nstep_between_samples;
    const bool atomic = run.scatter == SFX_SCATTER_ATOMIC;
    const size_t rec_stride = static_cast<size_t>(nstacey) * h.npe * h.ncomp;
    const int sig_top = (nstep - 1) / nsb;  // sample index 0 is the latest sampled step
    // a regular step (neither the first nor a checkpoint), all step-dependent
    // values read from the device counter, which counts down
    auto enqueue_regular = [&] {
      launch_sources(stream, 2, h.npe, 2 * h.nadj, 0, adj_order.p, adj_slot.p, gather.p, adj_array.p, adj_stf.p, adj,
                     d_step.p);
      launch_sources(stream, 2, h.npe, h.nsrc, 0, src_order.p, src_ispec.p, gather.p, src_array.p, stf.p, bwd,
                     d_step.p);
      for (size_t c = 0; c + 1 < h.color_offsets.size(); c++) {
        const int c0 = h.color_offsets[c], cnt = h.color_offsets[c + 1] - c0;
        launch_stiffness_pair_2d(stream, c0, cnt, atomic, gather.p, xix.p, xiz.p, gammax.p, gammaz.p, jac.p, kappa.p,
                                 mu.p, q, adj, bwd, 1, rec_index.p, record.p, 1, dt, h.nspec, Kmu, Kkappa, d_step.p,
                                 static_cast<long long>(rec_stride));
      }
      boundary(adj);
      launch_seismograms(stream, 2, h.npe, h.nrec, run.nseismo_types, types.p, 0, rec_ispec.p, gather.p, rec_h.p, bwd,
                         back.dto2, 1, seismo.p, PointGeom2D{xix.p, xiz.p, gammax.p, gammaz.p, kappa.p, mu.p}, q,
                         d_step.p, nsb, sig_top);
      launch_update_pair(stream, h.ncomp, adj, bwd, nc, 1, racc.p);
      launch_step_tick(stream, d_step.p, -1);
    };
    const bool graphs = graph_2d();
    if (graphs) set_step(nstep - 1);
    Throttle throttle(stream);
    int isig = 0;
    for (int istep = nstep - 1; istep >= 0; istep--) {
      const bool first = istep == nstep - 1;
      const bool ckpt = is_checkpoint(istep);
      const bool sample = istep % nsb == 0;
      if (graphs && !first && !ckpt) {
        if (!g_adj) g_adj = capture(enqueue_regular);
        SFX_HIP(hipGraphLaunch(g_adj, stream));
        throttle.tick();
        if (sample) isig++;
        continue;
      }
      launch_sources(stream, 2, h.npe, 2 * h.nadj, istep, adj_order.p, adj_slot.p, gather.p, adj_array.p, adj_stf.p,
                     adj);
      if (!first) sources(bwd, istep);
      const float *rec_i = record.p + static_cast<size_t>(istep) * rec_stride;
      for (size_t c = 0; c + 1 < h.color_offsets.size(); c++) {
        const int c0 = h.color_offsets[c], cnt = h.color_offsets[c + 1] - c0;
        launch_stiffness_pair_2d(stream, c0, cnt, atomic, gather.p, xix.p, xiz.p, gammax.p, gammaz.p, jac.p, kappa.p,
                                 mu.p, q, adj, bwd, first ? 0 : 1, rec_index.p, rec_i, ckpt ? 0 : 1, dt, h.nspec, Kmu,
                                 Kkappa);
      }
      boundary(adj);
      if (ckpt) {
        launch_divide_corrector(stream, h.ncomp, adj, nc, 1, 1);
        launch_divide_corrector(stream, h.ncomp, bwd, back, 1, 1);
        // after istep the backward state is the forward state after step
        // istep: reset it to bound the reconstruction's drift
        const size_t n3 = 3 * stride;
        SFX_HIP(hipMemcpyAsync(fields.p, snapshots.p + static_cast<size_t>(istep / snapshot_every) * n3,
                               n3 * sizeof(float), hipMemcpyDeviceToDevice, stream));
        launch_frechet_2d(stream, h.nspec, gather.p, xix.p, xiz.p, gammax.p, gammaz.p, kappa.p, mu.p, rho.p, q, adj,
                          bwd, dt, kacc.p);
        launch_rho_product(stream, h.ncomp, adj, bwd, dt, racc.p);
        if (sample) seismograms(bwd, isig++, 0);
        launch_predictor(stream, h.ncomp, adj, nc);
        launch_predictor(stream, h.ncomp, bwd, back);
      } else {
        // the first backward state is already corrected; later ones are
        // sampled as they will be after the backward corrector (-dt / 2)
        if (sample) seismograms(bwd, isig++, first ? 0 : 1, back.dto2);
        launch_update_pair(stream, h.ncomp, adj, bwd, nc, first ? 0 : 1, racc.p);
      }
      if (graphs) launch_step_tick(stream, d_step.p, -1);
      throttle.tick();
    }
    launch_frechet_finish(stream, static_cast<int>(npts), gather.p, kappa.p, mu.p, rho.p, racc.p, Kmu, Kkappa, kern.p);
    kern_dirty = false;
  }
};

namespace sfx {
const HostAssembly &engine_host(const sfx_engine *e) { return e->h; }
const sfx_run_desc &engine_run(const sfx_engine *e) { return e->run; }
}  // namespace sfx

namespace {

void check_engine(const sfx_engine *e) {
  if (!e) throw std::invalid_argument("null engine");
}

hipStream_t stream_of(const sfx_engine *e) { return e ? e->stream : nullptr; }

void check_device(const sfx_engine *e) {
  check_engine(e);
  if (e->host_only) throw std::invalid_argument("host-only engine (sfx_engine_create_host_*): no device state");
}

template <class Build> int create(Build &&build, const sfx_run_desc *run, sfx_engine **out, bool host_only) {
  return guarded([&] {
    if (!run || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    // The 3-D block path stays off on the device.  Its one GPU memory fault
    // has a cause shown on the host build (threads 200-255 of
    // block_step_3d_kernel read the quadrature rows past the kernel-argument
    // struct, fixed since), but the fixed kernel has not been run on a GPU.
    // Its host assembly (block layout) is still built for host-only engines.
    if (run->scatter == SFX_SCATTER_BLOCK && !host_only)
      throw std::invalid_argument("SFX_SCATTER_BLOCK is switched off on the device; use SFX_SCATTER_GATHER");
    auto e = std::make_unique<sfx_engine>();
    e->run = *run;
    e->h = build();
    // the element coloring and the sample count are host data: host-only
    // engines carry them too
    color_elements(e->h, run->scatter);
    e->nsig = (run->nstep + run->nstep_between_samples - 1) / run->nstep_between_samples;
    e->host_only = host_only;
    if (!host_only) e->upload();
    *out = e.release();
  });
}

template <class Mesh, class Assemble>
int create_engine(const Mesh *mesh, int nsources, const sfx_source *sources, int nreceivers,
                  const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out, bool host_only,
                  Assemble &&assemble) {
  return create(
      [&] {
        if (!mesh) throw std::invalid_argument("null mesh");
        if (nsources < 0 || nreceivers < 0) throw std::invalid_argument("negative count");
        if ((nsources && !sources) || (nreceivers && !receivers)) throw std::invalid_argument("null array");
        return assemble(*mesh, nsources, sources, nreceivers, receivers, *run);
      },
      run, out, host_only);
}

}  // namespace

extern "C" {

int sfx_version(void) { return 100; }

const char *sfx_last_error(void) { return last_error().c_str(); }

int sfx_engine_create_2d(const sfx_mesh2d *mesh, int nsources, const sfx_source *sources, int nreceivers,
                         const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out) {
  return create_engine(mesh, nsources, sources, nreceivers, receivers, run, out, false, build_assembly_2d);
}

int sfx_engine_create_3d(const sfx_mesh3d *mesh, int nsources, const sfx_source *sources, int nreceivers,
                         const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out) {
  return create_engine(mesh, nsources, sources, nreceivers, receivers, run, out, false, build_assembly_3d);
}

int sfx_engine_create_3d_db(const sfx_mesh3d_db *mesh, int nsources, const sfx_source *sources, int nreceivers,
                            const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out) {
  return create_engine(mesh, nsources, sources, nreceivers, receivers, run, out, false, build_assembly_3d_db);
}

int sfx_engine_create_host_3d_db(const sfx_mesh3d_db *mesh, int nsources, const sfx_source *sources,
                                 int nreceivers, const sfx_receiver *receivers, const sfx_run_desc *run,
                                 sfx_engine **out) {
  return create_engine(mesh, nsources, sources, nreceivers, receivers, run, out, true, build_assembly_3d_db);
}

int sfx_engine_create_host_2d(const sfx_mesh2d *mesh, int nsources, const sfx_source *sources, int nreceivers,
                              const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out) {
  return create_engine(mesh, nsources, sources, nreceivers, receivers, run, out, true, build_assembly_2d);
}

int sfx_engine_create_host_3d(const sfx_mesh3d *mesh, int nsources, const sfx_source *sources, int nreceivers,
                              const sfx_receiver *receivers, const sfx_run_desc *run, sfx_engine **out) {
  return create_engine(mesh, nsources, sources, nreceivers, receivers, run, out, true, build_assembly_3d);
}

void sfx_engine_destroy(sfx_engine *engine) {
  if (!engine) return;
  if (engine->stream) (void)hipStreamSynchronize(engine->stream);
  delete engine;
}

int sfx_get_sizes(const sfx_engine *e, int64_t *s) {
  return guarded([&] {
    check_engine(e);
    if (!s) throw std::invalid_argument("null output");
    s[0] = e->h.dim;
    s[1] = e->h.nspec;
    s[2] = e->h.nglob;
    s[3] = 5;
    s[4] = e->nsig;
    s[5] = e->h.nrec;
    s[6] = e->h.nsrc;
  });
}

int sfx_get_layout(const sfx_engine *e, int64_t *s) {
  return guarded([&] {
    check_engine(e);
    if (!s) throw std::invalid_argument("null output");
    s[0] = e->h.dim == 2 ? e->geom2d_form() : (e->h.general3d ? GEOM2D_POINT : GEOM2D_RECORD);
    s[1] = e->element_properties().empty() ? 0 : 1;
    s[2] = e->run.scatter;
    s[3] = e->gather2d() ? 1 : 0;
  });
}

int sfx_get_media_sizes(const sfx_engine *e, int64_t *s) {
  return guarded([&] {
    check_engine(e);
    if (!s) throw std::invalid_argument("null output");
    s[0] = e->h.nspec_el;
    s[1] = e->h.nspec - e->h.nspec_el;
    s[2] = e->h.nglob_el;
    s[3] = e->h.nglob_ac;
  });
}

int sfx_get_field_index_acoustic(const sfx_engine *e, int32_t *out) {
  return guarded([&] {
    check_engine(e);
    if (!out) throw std::invalid_argument("null output");
    if (e->h.field_index_ac.empty())
      std::fill(out, out + e->h.nglob, -1);
    else
      std::memcpy(out, e->h.field_index_ac.data(), e->h.field_index_ac.size() * sizeof(int32_t));
  });
}

int sfx_get_t0(const sfx_engine *e, double *t0) {
  return guarded([&] {
    check_engine(e);
    if (!t0) throw std::invalid_argument("null output");
    *t0 = e->h.t0;
  });
}

int sfx_get_index_mapping(const sfx_engine *e, int32_t *out) {
  return guarded([&] {
    check_engine(e);
    if (!out) throw std::invalid_argument("null output");
    std::memcpy(out, e->h.index_mapping.data(), e->h.index_mapping.size() * sizeof(int32_t));
  });
}

int sfx_get_field_index(const sfx_engine *e, int32_t *out) {
  return guarded([&] {
    check_engine(e);
    if (!out) throw std::invalid_argument("null output");
    std::memcpy(out, e->h.field_index.data(), e->h.field_index.size() * sizeof(int32_t));
  });
}

int sfx_get_compute_to_mesh(const sfx_engine *e, int32_t *out) {
  return guarded([&] {
    check_engine(e);
    if (!out) throw std::invalid_argument("null output");
    std::memcpy(out, e->h.compute_to_mesh.data(), e->h.compute_to_mesh.size() * sizeof(int32_t));
  });
}

int sfx_get_locations(const sfx_engine *e, int which, int32_t *ispec, double *local) {
  return guarded([&] {
    check_engine(e);
    if (which != 0 && which != 1) throw std::invalid_argument("which must be 0 or 1");
    const auto &is = which == 0 ? e->h.src_ispec : e->h.rec_ispec;
    const auto &lc = which == 0 ? e->h.src_local : e->h.rec_local;
    if (ispec) std::memcpy(ispec, is.data(), is.size() * sizeof(int32_t));
    if (local) std::memcpy(local, lc.data(), lc.size() * sizeof(double));
  });
}

int sfx_initialize(sfx_engine *e) {
  return guarded([&] {
    check_device(e);
    const size_t stride = static_cast<size_t>(e->h.ncomp) * e->h.nglob_pad;
    launch_fill(e->stream, e->fields.p, 3 * stride, 0.0f);
    if (e->h.coupled()) launch_fill(e->stream, e->ac_fields.p, 3 * static_cast<size_t>(e->h.nglob_ac_pad), 0.0f);
    if (e->seismo.n) launch_fill(e->stream, e->seismo.p, e->seismo.n, 0.0f);
    e->a_stale = false;
    SFX_HIP(hipGetLastError());
    e->initialized = true;
  }, stream_of(e));
}

int sfx_predictor(sfx_engine *e) {
  int dofs = 0;
  const int rc = guarded([&] {
    check_device(e);
    e->require_init();
    e->refresh_acceleration();
    launch_predictor(e->stream, e->h.ncomp, e->f(), e->nc, e->gather2d());
    SFX_HIP(hipGetLastError());
    dofs = e->h.nglob_el * e->h.ncomp;
  }, stream_of(e));
  return rc == SFX_OK ? dofs : rc;
}

int sfx_update_wavefields(sfx_engine *e, int istep) {
  int nelem = 0;
  const int rc = guarded([&] {
    check_device(e);
    e->require_init();
    e->update_forward(istep);
    SFX_HIP(hipGetLastError());
    nelem = e->h.nspec_el;
  }, stream_of(e));
  return rc == SFX_OK ? nelem : rc;
}

int sfx_corrector(sfx_engine *e) {
  int dofs = 0;
  const int rc = guarded([&] {
    check_device(e);
    e->require_init();
    e->refresh_acceleration();
    launch_divide_corrector(e->stream, e->h.ncomp, e->f(), e->nc, 0, 1, e->gather2d());
    SFX_HIP(hipGetLastError());
    dofs = e->h.nglob_el * e->h.ncomp;
  }, stream_of(e));
  return rc == SFX_OK ? dofs : rc;
}

int sfx_compute_seismograms(sfx_engine *e, int isig) {
  return guarded([&] {
    check_device(e);
    e->require_init();
    e->refresh_acceleration();
    e->seismograms(e->f(), isig, 0);
    SFX_HIP(hipGetLastError());
  }, stream_of(e));
}

int sfx_initialize_wavefield(sfx_engine *e, int wavefield) {
  return guarded([&] {
    check_device(e);
    if (wavefield == SFX_WAVEFIELD_FORWARD) {
      const int rc = sfx_initialize(e);
      if (rc != SFX_OK) throw std::runtime_error(last_error());
      return;
    }
    e->require_adjoint_capable();
    Fields F = e->wavefield(wavefield);
    const size_t stride = static_cast<size_t>(e->h.ncomp) * e->h.nglob_pad;
    launch_fill(e->stream, F.u, 3 * stride, 0.0f);
    if (wavefield == SFX_WAVEFIELD_ADJOINT) e->frechet_zero();
    SFX_HIP(hipGetLastError());
    e->initialized = true;
  }, stream_of(e));
}

int sfx_predictor_phase(sfx_engine *e, int medium, int wavefield, int direction) {
  int dofs = 0;
  const int rc = guarded([&] {
    check_device(e);
    e->require_init();
    sfx_engine::check_medium(medium);
    const NewmarkConst k = e->newmark(direction);
    dofs = e->medium_dofs(medium);
    if (!dofs) return;
    int ncomp = 0;
    Fields F = e->medium_wavefield(medium, wavefield, ncomp);
    if (wavefield == SFX_WAVEFIELD_FORWARD) e->refresh_acceleration();
    launch_predictor(e->stream, ncomp, F, k, e->gather2d());
    SFX_HIP(hipGetLastError());
  }, stream_of(e));
  return rc == SFX_OK ? dofs : rc;
}

int sfx_corrector_phase(sfx_engine *e, int medium, int wavefield, int direction) {
  int dofs = 0;
  const int rc = guarded([&] {
    check_device(e);
    e->require_init();
    sfx_engine::check_medium(medium);
    const NewmarkConst k = e->newmark(direction);
    dofs = e->medium_dofs(medium);
    if (!dofs) return;
    int ncomp = 0;
    Fields F = e->medium_wavefield(medium, wavefield, ncomp);
    if (wavefield == SFX_WAVEFIELD_FORWARD) e->refresh_acceleration();
    launch_divide_corrector(e->stream, ncomp, F, k, 0, 1, e->gather2d());
    SFX_HIP(hipGetLastError());
  }, stream_of(e));
  return rc == SFX_OK ? dofs : rc;
}

int sfx_update_wavefields_medium(sfx_engine *e, int medium, int wavefield, int istep) {
  int nelem = 0;
  const int rc = guarded([&] {
    check_device(e);
    e->require_init();
    sfx_engine::check_medium(medium);
    nelem = e->medium_elements(medium);
    if (!nelem) return;
    if (medium == SFX_MEDIUM_ACOUSTIC) {
      if (wavefield != SFX_WAVEFIELD_FORWARD) throw std::invalid_argument("acoustic media have only the forward wavefield");
      e->update_acoustic(istep);
    } else if (wavefield == SFX_WAVEFIELD_FORWARD)
      e->update_forward(istep);
    else if (wavefield == SFX_WAVEFIELD_ADJOINT)
      e->update_adjoint(istep);
    else if (wavefield == SFX_WAVEFIELD_BACKWARD)
      e->update_backward(istep);
    else
      throw std::invalid_argument("the buffer wavefield is not stepped");
    SFX_HIP(hipGetLastError());
  }, stream_of(e));
  return rc == SFX_OK ? nelem : rc;
}

int sfx_compute_seismograms_wavefield(sfx_engine *e, int wavefield, int isig) {
  return guarded([&] {
    check_device(e);
    e->require_init();
    if (wavefield == SFX_WAVEFIELD_FORWARD) e->refresh_acceleration();
    e->seismograms(e->wavefield(wavefield), isig, 0);
    SFX_HIP(hipGetLastError());
  }, stream_of(e));
}

int sfx_copy_wavefield(sfx_engine *e, int dst, int src) {
  return guarded([&] {
    check_device(e);
    if (dst == src) throw std::invalid_argument("source and destination wavefields are the same");
    if (src == SFX_WAVEFIELD_FORWARD) e->refresh_acceleration();
    Fields from = e->wavefield(src), to = e->wavefield(dst);
    const size_t n3 = 3 * static_cast<size_t>(e->h.ncomp) * e->h.nglob_pad;
    SFX_HIP(hipMemcpyAsync(to.u, from.u, n3 * sizeof(float), hipMemcpyDeviceToDevice, e->stream));
    if (dst == SFX_WAVEFIELD_FORWARD) e->a_stale = false;
  }, stream_of(e));
}

int sfx_compute_derivatives(sfx_engine *e, double dt) {
  return guarded([&] {
    check_device(e);
    e->require_init();
    e->require_adjoint_capable();
    if (!e->kacc.n || !e->racc.n) e->frechet_zero();
    Fields adj = e->wavefield(SFX_WAVEFIELD_ADJOINT), bwd = e->wavefield(SFX_WAVEFIELD_BACKWARD);
    const float fdt = static_cast<float>(dt);
    launch_frechet_2d(e->stream, e->h.nspec, e->gather.p, e->xix.p, e->xiz.p, e->gammax.p, e->gammaz.p, e->kappa.p,
                      e->mu.p, e->rho.p, e->q, adj, bwd, fdt, e->kacc.p);
    launch_rho_product(e->stream, e->h.ncomp, adj, bwd, fdt, e->racc.p);
    e->kern_dirty = true;
    SFX_HIP(hipGetLastError());
  }, stream_of(e));
}

int sfx_step(sfx_engine *e, int istep) {
  return guarded([&] {
    check_device(e);
    e->require_init();
    e->own_forces();
    e->step(istep, nullptr, true);
    SFX_HIP(hipGetLastError());
  }, stream_of(e));
}

int sfx_run(sfx_engine *e, int first, int count) {
  return guarded([&] {
    check_device(e);
    e->require_init();
    if (first < 0 || count < 0 || first + count > e->run.nstep) throw std::invalid_argument("step range out of bounds");
    e->own_forces();
    if (e->graph_2d() && count > 1) {
      // 2-D: one captured graph per step; only the whole-step time is taken
      e->set_step(first);
      if (!e->g_fwd) e->g_fwd = e->capture([&] { e->enqueue_forward_2d(false); });
      while (e->ev.size() < 2) {
        hipEvent_t ev;
        SFX_HIP(hipEventCreate(&ev));
        e->ev.push_back(ev);
      }
      SFX_HIP(hipEventRecord(e->ev[0], e->stream));
      Throttle throttle(e->stream);
      for (int i = 0; i < count; i++) {
        SFX_HIP(hipGraphLaunch(e->g_fwd, e->stream));
        throttle.tick();
      }
      SFX_HIP(hipEventRecord(e->ev[1], e->stream));
      if (e->gather2d()) e->a_stale = true;
      SFX_HIP(hipStreamSynchronize(e->stream));
      float all = 0.0f;
      SFX_HIP(hipEventElapsedTime(&all, e->ev[0], e->ev[1]));
      e->timings[0] = e->timings[1] = e->timings[2] = 0.0;
      e->timings[3] = all / count;
      return;
    }
    // Every step records 5 events on the engine stream; they are read after
    // the last step (the host only waits to keep the queue bounded).
    const size_t need = 5 * static_cast<size_t>(count);
    while (e->ev.size() < need) {
      hipEvent_t ev;
      SFX_HIP(hipEventCreate(&ev));
      e->ev.push_back(ev);
    }
    Throttle throttle(e->stream);
    for (int i = 0; i < count; i++) {
      e->step(first + i, &e->ev[5 * static_cast<size_t>(i)], i == count - 1);
      throttle.tick();
    }
    SFX_HIP(hipGetLastError());
    SFX_HIP(hipStreamSynchronize(e->stream));
    double acc[4] = {0, 0, 0, 0};
    for (int i = 0; i < count; i++) {
      hipEvent_t *v = &e->ev[5 * static_cast<size_t>(i)];
      float stiff, upd, src, bnd, all;
      SFX_HIP(hipEventElapsedTime(&src, v[0], v[1]));
      SFX_HIP(hipEventElapsedTime(&stiff, v[1], v[2]));
      SFX_HIP(hipEventElapsedTime(&bnd, v[2], v[3]));
      SFX_HIP(hipEventElapsedTime(&upd, v[3], v[4]));
      SFX_HIP(hipEventElapsedTime(&all, v[0], v[4]));
      acc[0] += stiff;
      acc[1] += upd;
      acc[2] += src + bnd;
      acc[3] += all;
    }
    for (int k = 0; k < 4; k++) e->timings[k] = count ? acc[k] / count : 0.0;
  }, stream_of(e));
}

int sfx_run_noalloc(sfx_engine *e, int first, int count) {
  return guarded([&] {
    check_device(e);
    e->require_init();
    if (!e->gather_path()) throw std::invalid_argument("sfx_run_noalloc needs a 3-D engine with SFX_SCATTER_GATHER");
    if (first < 0 || count < 0 || first + count > e->run.nstep) throw std::invalid_argument("step range out of bounds");
    if (e->h.nspec > SFX_NOALLOC_MAX_ELEMENTS)
      throw std::invalid_argument("sfx_run_noalloc: " + std::to_string(e->h.nspec) + " elements exceed the scratch limit of " +
                                  std::to_string(SFX_NOALLOC_MAX_ELEMENTS) + " (SFX_NOALLOC_MAX_ELEMENTS)");
    if (count == 0) return;
    float *const sc = scratch_on(e->device);
    // an acceleration left behind by sfx_run is assembled from the engine's
    // own buffer first (it exists whenever one is left behind)
    e->refresh_acceleration();
    // every step writes each element's forces before any kernel reads them,
    // so the shared scratch needs no clearing between calls
    e->re_p = sc;
    try {
      for (int i = 0; i < count; i++) {
        e->step(first + i, nullptr, i == count - 1);
        // bound the queued work without creating events: wait every 256 steps
        if ((i + 1) % 256 == 0) SFX_HIP(hipStreamSynchronize(e->stream));
      }
    } catch (...) {
      if (e->a_stale) launch_assemble_3d(e->stream, e->lat, sc, e->f(), 1);
      e->a_stale = false;
      e->re_p = e->re.p;
      throw;
    }
    // assemble the stored acceleration from the scratch now, while it holds
    // this call's last forces
    launch_assemble_3d(e->stream, e->lat, sc, e->f(), 1);
    e->a_stale = false;
    e->re_p = e->re.p;
    SFX_HIP(hipGetLastError());
    SFX_HIP(hipStreamSynchronize(e->stream));
  }, stream_of(e));
}

int sfx_synchronize(sfx_engine *e) {
  return guarded([&] {
    check_device(e);
    SFX_HIP(hipStreamSynchronize(e->stream));
  }, stream_of(e));
}

int sfx_get_seismograms(const sfx_engine *e, float *out) {
  return guarded([&] {
    check_device(e);
    if (!out) throw std::invalid_argument("null output");
    SFX_HIP(hipStreamSynchronize(e->stream));
    if (e->seismo.n) SFX_HIP(hipMemcpy(out, e->seismo.p, e->seismo.n * sizeof(float), hipMemcpyDeviceToHost));
  }, stream_of(e));
}

int sfx_get_wavefield(const sfx_engine *ce, int wavefield, int which, float *out) {
  return guarded([&] {
    check_device(ce);
    if (which < 0 || which > 3) throw std::invalid_argument("which must be 0..3");
    if (!out) throw std::invalid_argument("null output");
    // the stored acceleration is refreshed lazily after fused gather steps
    sfx_engine *e = const_cast<sfx_engine *>(ce);
    if (wavefield == SFX_WAVEFIELD_FORWARD && which == 2) e->refresh_acceleration();
    Fields F = e->wavefield(wavefield);
    const float *arrays[4] = {F.u, F.v, F.a, F.minv};
    SFX_HIP(hipStreamSynchronize(e->stream));
    const size_t ng = e->h.nglob_el;
    std::vector<float> buf(e->h.nglob_pad);
    for (int c = 0; c < e->h.ncomp; c++) {
      SFX_HIP(hipMemcpy(buf.data(), arrays[which] + static_cast<size_t>(c) * e->h.nglob_pad,
                        e->h.nglob_pad * sizeof(float), hipMemcpyDeviceToHost));
      float *o = out + static_cast<size_t>(c) * ng;
      if (e->h.lattice_of.empty())
        std::memcpy(o, buf.data(), ng * sizeof(float));
      else
        for (size_t g = 0; g < ng; g++) o[g] = buf[e->h.lattice_of[g]];
    }
  }, stream_of(ce));
}

int sfx_get_field(const sfx_engine *e, int which, float *out) {
  return sfx_get_wavefield(e, SFX_WAVEFIELD_FORWARD, which, out);
}

int sfx_set_wavefield(sfx_engine *e, int wavefield, int which, const float *in) {
  return guarded([&] {
    check_device(e);
    if (which < 0 || which > 2) throw std::invalid_argument("which must be 0..2");
    if (!in) throw std::invalid_argument("null input");
    Fields F = e->wavefield(wavefield);
    float *arrays[3] = {F.u, F.v, F.a};
    SFX_HIP(hipStreamSynchronize(e->stream));
    const size_t ng = e->h.nglob_el;
    std::vector<float> buf(e->h.nglob_pad, 0.0f);
    for (int c = 0; c < e->h.ncomp; c++) {
      const float *src = in + static_cast<size_t>(c) * ng;
      if (e->h.lattice_of.empty())
        std::memcpy(buf.data(), src, ng * sizeof(float));
      else
        for (size_t g = 0; g < ng; g++) buf[e->h.lattice_of[g]] = src[g];
      SFX_HIP(hipMemcpy(arrays[which] + static_cast<size_t>(c) * e->h.nglob_pad, buf.data(),
                        e->h.nglob_pad * sizeof(float), hipMemcpyHostToDevice));
    }
    if (wavefield == SFX_WAVEFIELD_FORWARD && which == 2) e->a_stale = false;
  }, stream_of(e));
}

int sfx_set_field(sfx_engine *e, int which, const float *in) {
  return sfx_set_wavefield(e, SFX_WAVEFIELD_FORWARD, which, in);
}

int sfx_get_medium_field(const sfx_engine *ce, int medium, int which, float *out) {
  if (medium != SFX_MEDIUM_ACOUSTIC) return sfx_get_wavefield(ce, SFX_WAVEFIELD_FORWARD, which, out);
  return guarded([&] {
    check_device(ce);
    if (which < 0 || which > 3) throw std::invalid_argument("which must be 0..3");
    if (!out) throw std::invalid_argument("null output");
    if (!ce->h.coupled()) throw std::invalid_argument("the mesh has no acoustic element");
    const Fields F = ce->fa();
    const float *arrays[4] = {F.u, F.v, F.a, F.minv};
    SFX_HIP(hipStreamSynchronize(ce->stream));
    if (ce->h.nglob_ac)
      SFX_HIP(hipMemcpy(out, arrays[which], ce->h.nglob_ac * sizeof(float), hipMemcpyDeviceToHost));
  }, stream_of(ce));
}

int sfx_set_medium_field(sfx_engine *e, int medium, int which, const float *in) {
  if (medium != SFX_MEDIUM_ACOUSTIC) return sfx_set_wavefield(e, SFX_WAVEFIELD_FORWARD, which, in);
  return guarded([&] {
    check_device(e);
    if (which < 0 || which > 2) throw std::invalid_argument("which must be 0..2");
    if (!in) throw std::invalid_argument("null input");
    if (!e->h.coupled()) throw std::invalid_argument("the mesh has no acoustic element");
    const Fields F = e->fa();
    float *arrays[3] = {F.u, F.v, F.a};
    SFX_HIP(hipStreamSynchronize(e->stream));
    if (e->h.nglob_ac)
      SFX_HIP(hipMemcpy(arrays[which], in, e->h.nglob_ac * sizeof(float), hipMemcpyHostToDevice));
  }, stream_of(e));
}

int sfx_set_properties(sfx_engine *e, const float *rho, const float *kappa, const float *mu) {
  return guarded([&] {
    check_device(e);
    SFX_HIP(hipStreamSynchronize(e->stream));
    set_properties_2d(e->h, rho, kappa, mu, e->run.dt);
    hipStream_t s = e->stream;
    e->kappa.upload(e->permuted(e->h.kappa), s);
    e->mu.upload(e->permuted(e->h.mu), s);
    e->upload_element_properties(s);
    e->rho.upload(e->permuted(e->h.rho), s);
    e->bnd_rvp.upload(e->h.bnd_rho_vp, s);
    e->bnd_rvs.upload(e->h.bnd_rho_vs, s);
    const size_t stride = static_cast<size_t>(e->h.ncomp) * e->h.nglob_pad;
    SFX_HIP(hipMemcpyAsync(e->fields.p + 3 * stride, e->h.mass_inverse.data(), stride * sizeof(float),
                           hipMemcpyHostToDevice, s));
    if (e->gather2d()) e->upload_gather_2d(s);  // the sorted Stacey entries carry the impedances
    SFX_HIP(hipStreamSynchronize(s));
    e->recorded = false;
    e->invalidate_graphs();
  }, stream_of(e));
}

int sfx_forward_record(sfx_engine *e, int snapshot_every) {
  return guarded([&] {
    check_device(e);
    e->require_init();
    e->require_adjoint_capable();
    if (snapshot_every < 0) throw std::invalid_argument("snapshot_every must be >= 0");
    const size_t need = static_cast<size_t>(e->run.nstep) * e->nstacey * e->h.npe * e->h.ncomp;
    if (e->record.n != need) e->record.alloc(need);
    e->snapshot_every = snapshot_every;
    const int nsnap = snapshot_every > 0 && e->run.nstep >= 2 ? (e->run.nstep - 2) / snapshot_every + 1 : 0;
    const size_t nsnap_f = static_cast<size_t>(nsnap) * 3 * e->h.ncomp * e->h.nglob_pad;
    if (e->snapshots.n != nsnap_f) e->snapshots.alloc(nsnap_f);
    e->recorded = false;
    e->recording = true;
    try {
      const bool graphs = e->graph_2d();
      if (graphs) e->set_step(0);
      Throttle throttle(e->stream);
      for (int istep = 0; istep < e->run.nstep; istep++) {
        const bool last = istep == e->run.nstep - 1;
        if (graphs && !last && !e->is_checkpoint(istep)) {
          if (!e->g_rec) e->g_rec = e->capture([&] { e->enqueue_forward_2d(true); });
          SFX_HIP(hipGraphLaunch(e->g_rec, e->stream));
        } else {
          e->step(istep, nullptr, last);
          if (graphs) launch_step_tick(e->stream, e->d_step.p, 1);
        }
        throttle.tick();
      }
    } catch (...) {
      e->recording = false;
      throw;
    }
    e->recording = false;
    SFX_HIP(hipGetLastError());
    SFX_HIP(hipStreamSynchronize(e->stream));
    e->recorded = true;
  }, stream_of(e));
}

int sfx_set_adjoint_sources(sfx_engine *e, int nadj, const sfx_receiver *positions, const float *stf) {
  return guarded([&] {
    check_device(e);
    place_adjoint_sources_2d(e->h, nadj, positions, stf, e->run.nstep);
    hipStream_t s = e->stream;
    e->adj_order.upload(e->h.adj_order, s);
    e->adj_slot.upload(e->slots_of(e->h.adj_ispec), s);
    e->adj_array.upload(e->h.adj_array, s);
    e->adj_stf.upload(e->h.adj_stf, s);
    SFX_HIP(hipStreamSynchronize(s));
    e->adj_ready = true;
    e->invalidate_graphs();
  }, stream_of(e));
}

int sfx_adjoint_run(sfx_engine *e) {
  return guarded([&] {
    check_device(e);
    if (e->h.dim != 2) throw std::invalid_argument("adjoint runs are implemented for 2-D meshes");
    if (!e->recorded) throw std::logic_error("sfx_forward_record must complete before sfx_adjoint_run");
    if (!e->adj_ready) throw std::logic_error("sfx_set_adjoint_sources must be called before sfx_adjoint_run");
    e->recorded = false;  // the backward field consumes the forward state
    e->adjoint_run();
    SFX_HIP(hipGetLastError());
    SFX_HIP(hipStreamSynchronize(e->stream));
  }, stream_of(e));
}

int sfx_get_point_coordinates(const sfx_engine *e, float *out) {
  return guarded([&] {
    check_engine(e);
    if (!out) throw std::invalid_argument("null output");
    if (e->h.dim != 2) throw std::invalid_argument("point coordinates are implemented for 2-D meshes");
    const size_t n = static_cast<size_t>(e->h.nspec) * e->h.npe;
    for (size_t p = 0; p < n; p++) {
      out[2 * p] = e->h.px2d[p];
      out[2 * p + 1] = e->h.pz2d[p];
    }
  });
}

int sfx_get_stacey_record(const sfx_engine *e, float *out) {
  return guarded([&] {
    check_device(e);
    if (!out) throw std::invalid_argument("null output");
    const size_t per_step = static_cast<size_t>(e->nstacey) * e->h.npe * e->h.ncomp;
    if (e->h.dim != 2 || e->record.n != static_cast<size_t>(e->run.nstep) * per_step)
      throw std::logic_error("no Stacey record: run sfx_forward_record first");
    SFX_HIP(hipStreamSynchronize(e->stream));
    std::vector<float> buf(e->record.n);
    if (!buf.empty())
      SFX_HIP(hipMemcpy(buf.data(), e->record.p, buf.size() * sizeof(float), hipMemcpyDeviceToHost));
    // device (nstep, nstacey, 25, 2) -> boundary_medium_container values
    // (nstacey, ngllz, ngllx, nstep, 2) LayoutLeft: element fastest
    const size_t ns = e->nstacey, nt = e->run.nstep;
    for (size_t t = 0; t < nt; t++)
      for (size_t s = 0; s < ns; s++)
        for (size_t p = 0; p < 25; p++)
          for (size_t c = 0; c < 2; c++) out[s + ns * (p % 5 * 5 + p / 5 + 25 * (t + nt * c))] = buf[((t * ns + s) * 25 + p) * 2 + c];
  }, stream_of(e));
}

int sfx_set_stacey_record(sfx_engine *e, const float *in) {
  return guarded([&] {
    check_device(e);
    if (!in) throw std::invalid_argument("null input");
    e->require_adjoint_capable();
    const size_t ns = e->nstacey, nt = e->run.nstep;
    std::vector<float> buf(nt * ns * 25 * 2);
    for (size_t t = 0; t < nt; t++)
      for (size_t s = 0; s < ns; s++)
        for (size_t p = 0; p < 25; p++)
          for (size_t c = 0; c < 2; c++) buf[((t * ns + s) * 25 + p) * 2 + c] = in[s + ns * (p % 5 * 5 + p / 5 + 25 * (t + nt * c))];
    SFX_HIP(hipStreamSynchronize(e->stream));
    if (e->record.n != buf.size()) e->record.alloc(buf.size());
    if (!buf.empty())
      SFX_HIP(hipMemcpy(e->record.p, buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice));
    e->snapshot_every = 0;
    e->recorded = true;
    e->invalidate_graphs();
  }, stream_of(e));
}

int sfx_get_point_weights(const sfx_engine *e, float *out) {
  return guarded([&] {
    check_device(e);
    if (!out) throw std::invalid_argument("null output");
    if (e->h.dim != 2) throw std::invalid_argument("point weights are implemented for 2-D meshes");
    const size_t npe = e->h.npe;
    for (int c = 0; c < e->h.nspec; c++)
      for (int iz = 0; iz < 5; iz++)
        for (int ix = 0; ix < 5; ix++) {
          const size_t p = static_cast<size_t>(c) * npe + iz * 5 + ix;
          out[static_cast<size_t>(e->h.compute_to_mesh[c]) * npe + iz * 5 + ix] = e->h.w[ix] * e->h.w[iz] * e->h.jac[p];
        }
  }, stream_of(e));
}

int sfx_get_kernels(const sfx_engine *e, float *out) {
  return guarded([&] {
    check_device(e);
    if (!out) throw std::invalid_argument("null output");
    if (!e->kern.n && !e->kern_dirty) throw std::logic_error("no kernels: run sfx_adjoint_run first");
    if (e->kern_dirty) const_cast<sfx_engine *>(e)->frechet_finish();
    SFX_HIP(hipStreamSynchronize(e->stream));
    std::vector<float> buf(e->kern.n);
    SFX_HIP(hipMemcpy(buf.data(), e->kern.p, buf.size() * sizeof(float), hipMemcpyDeviceToHost));
    const size_t npe = e->h.npe, nspec = e->h.nspec;
    for (int k = 0; k < 6; k++)
      for (size_t slot = 0; slot < nspec; slot++) {
        const size_t m = e->h.compute_to_mesh[e->h.perm[slot]];
        std::memcpy(out + (k * nspec + m) * npe, buf.data() + (k * nspec + slot) * npe, npe * sizeof(float));
      }
  }, stream_of(e));
}

int sfx_get_timings(const sfx_engine *e, double *out4) {
  return guarded([&] {
    check_device(e);
    if (!out4) throw std::invalid_argument("null output");
    for (int k = 0; k < 4; k++) out4[k] = e->timings[k];
  }, stream_of(e));
}

}  // extern "C"
