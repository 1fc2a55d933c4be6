// Host assembly of the 2-D and 3-D engines.  See assembly.hpp for the layout.
#include "assembly.hpp"
#include "lattice.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>

#include "quadrature.hpp"

namespace sfx {

namespace {

// Unsigned key that orders like the float (+0 and -0, which compare equal,
// share one key).
inline uint32_t float_key(float f) {
  if (f == 0.0f) f = 0.0f;
  uint32_t u;
  std::memcpy(&u, &f, sizeof u);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// Indices [0, n) ordered by (x, z) as float comparisons order them: LSD radix
// sort on the 64-bit key (key(x), key(z)), 11-bit digits, passes over a
// constant digit skipped.  Points with equal keys have equal coordinates, so
// their relative order cannot change the merge or the numbering below.
std::vector<uint32_t> sort_by_xz(const std::vector<float> &x, const std::vector<float> &z) {
  const size_t n = x.size();
  std::vector<uint64_t> key(n), key2(n);
  std::vector<uint32_t> idx(n), idx2(n);
  for (size_t i = 0; i < n; i++) {
    key[i] = (static_cast<uint64_t>(float_key(x[i])) << 32) | float_key(z[i]);
    idx[i] = static_cast<uint32_t>(i);
  }
  constexpr int BITS = 11, RADIX = 1 << BITS;
  std::vector<size_t> count(RADIX);
  for (int shift = 0; shift < 64; shift += BITS) {
    std::fill(count.begin(), count.end(), 0);
    for (size_t i = 0; i < n; i++) count[(key[i] >> shift) & (RADIX - 1)]++;
    if (count[(key[0] >> shift) & (RADIX - 1)] == n) continue;
    size_t sum = 0;
    for (auto &c : count) {
      const size_t t = c;
      c = sum;
      sum += t;
    }
    for (size_t i = 0; i < n; i++) {
      const size_t d = count[(key[i] >> shift) & (RADIX - 1)]++;
      key2[d] = key[i];
      idx2[d] = idx[i];
    }
    key.swap(key2);
    idx.swap(idx2);
  }
  return idx;
}

// Element boundary tags (src/mesh/dim2/tags/tags.cpp:5-47), combined as the
// tag container's += does (include/enumerations/boundary.hpp:50-90): a Stacey
// and an acoustic free-surface tag on one element make it composite.
enum BoundaryTag { TAG_NONE = 0, TAG_FREE_SURFACE = 1, TAG_STACEY = 2, TAG_COMPOSITE = 3 };
int add_tag(int cur, int t) {
  if (cur == TAG_NONE) return t;
  if (cur == t || cur == TAG_COMPOSITE) return cur;
  return TAG_COMPOSITE;
}

// Mesh-order element types -> compute order
// (src/compute/compute_mesh.cpp:240-279 with the ELEMENT_TYPES order of
// include/enumerations/material_definitions.hpp:76-103: elastic P-SV
// isotropic without boundary, with Stacey; acoustic isotropic without
// boundary, with a free surface, with Stacey, composite), each group in mesh
// order.  acoustic_mesh: 1 for the elements of acoustic materials.
void compute_ordering(const sfx_mesh2d &m, const std::vector<int> &acoustic_mesh, HostAssembly &a,
                      std::vector<int> &tag_mesh) {
  tag_mesh.assign(m.nspec, TAG_NONE);
  for (int i = 0; i < m.nabs; i++) {
    if (m.abs_ispec[i] < 0 || m.abs_ispec[i] >= m.nspec) throw std::invalid_argument("abs_ispec out of range");
    if (m.abs_edge[i] < 0 || m.abs_edge[i] > 3) throw std::invalid_argument("abs_edge out of range");
    tag_mesh[m.abs_ispec[i]] = add_tag(tag_mesh[m.abs_ispec[i]], TAG_STACEY);
  }
  for (int i = 0; i < m.nafs; i++) {
    const int e = m.afs_ispec[i];
    if (e < 0 || e >= m.nspec) throw std::invalid_argument("afs_ispec out of range");
    if (m.afs_type[i] < 0 || m.afs_type[i] > 7) throw std::invalid_argument("afs_type out of range");
    if (!acoustic_mesh[e]) throw std::invalid_argument("Acoustic free surface boundary is not an acoustic element");
    tag_mesh[e] = add_tag(tag_mesh[e], TAG_FREE_SURFACE);
  }
  const int groups[6][2] = {{0, TAG_NONE}, {0, TAG_STACEY}, {1, TAG_NONE}, {1, TAG_FREE_SURFACE}, {1, TAG_STACEY},
                            {1, TAG_COMPOSITE}};
  a.compute_to_mesh.clear();
  a.compute_to_mesh.reserve(m.nspec);
  for (const auto &grp : groups)
    for (int e = 0; e < m.nspec; e++)
      if (acoustic_mesh[e] == grp[0] && tag_mesh[e] == grp[1]) a.compute_to_mesh.push_back(e);
  a.nspec_none = 0;
  a.nspec_el = 0;
  for (int e = 0; e < m.nspec; e++) {
    a.nspec_el += acoustic_mesh[e] ? 0 : 1;
    a.nspec_none += (!acoustic_mesh[e] && tag_mesh[e] == TAG_NONE) ? 1 : 0;
  }
}

// Tolerance of the point merge, src/compute/compute_mesh.cpp:19-43: taken on
// consecutive groups of ngllxz *sorted* points, starting the maxima at the
// smallest positive float, exactly as the reference does.
float merge_tolerance(const std::vector<uint32_t> &order, const std::vector<float> &x,
                      const std::vector<float> &z, int nspec, int ngllxz) {
  float typ = std::numeric_limits<float>::max();
  for (int e = 0; e < nspec; e++) {
    float xmax = std::numeric_limits<float>::min(), xmin = std::numeric_limits<float>::max();
    float zmax = std::numeric_limits<float>::min(), zmin = std::numeric_limits<float>::max();
    const uint32_t *g = &order[static_cast<size_t>(e) * ngllxz];
    for (int k = 0; k < ngllxz; k++) {
      xmax = std::max(xmax, x[g[k]]);
      xmin = std::min(xmin, x[g[k]]);
      zmax = std::max(zmax, z[g[k]]);
      zmin = std::min(zmin, z[g[k]]);
    }
    typ = std::min(typ, xmax - xmin);
    typ = std::min(typ, zmax - zmin);
  }
  return 1e-6 * typ;
}

// Numbering, src/compute/compute_mesh.cpp:45-170: sort the points by (x, z),
// merge neighbours closer than the tolerance, then number by first visit in
// (chunk, iz, ix, element) order.
void number_points_2d(HostAssembly &a, const std::vector<float> &px, const std::vector<float> &pz, int chunk) {
  const int nspec = a.nspec, npe = 25;
  const size_t nloc = static_cast<size_t>(nspec) * npe;
  // visit order iloc -> local point p = e * 25 + iz * 5 + ix
  std::vector<uint32_t> visit(nloc);
  std::vector<float> vx(nloc), vz(nloc);
  size_t iloc = 0;
  for (int c0 = 0; c0 < nspec; c0 += chunk)
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++)
        for (int ie = 0; ie < chunk && c0 + ie < nspec; ie++) {
          const size_t p = static_cast<size_t>(c0 + ie) * npe + iz * NGLL + ix;
          visit[iloc] = static_cast<uint32_t>(p);
          vx[iloc] = px[p];
          vz[iloc] = pz[p];
          iloc++;
        }
  const std::vector<uint32_t> order = sort_by_xz(vx, vz);
  const float tol = merge_tolerance(order, vx, vz, nspec, npe);
  std::vector<int32_t> iglob(nloc);
  int ig = 0;
  iglob[order[0]] = 0;
  for (size_t i = 1; i < nloc; i++) {
    const uint32_t c = order[i], p = order[i - 1];
    if (std::abs(vx[c] - vx[p]) > tol || std::abs(vz[c] - vz[p]) > tol) ig++;
    iglob[c] = ig;
  }
  const int nglob = ig + 1;
  std::vector<int> first(nglob, -1);
  a.index_mapping.assign(nloc, -1);
  int inum = 0;
  for (size_t i = 0; i < nloc; i++) {
    int &f = first[iglob[i]];
    if (f < 0) f = inum++;
    a.index_mapping[visit[i]] = f;
  }
  if (inum != nglob) throw std::runtime_error("numbering: merged point count mismatch");
  a.nglob = nglob;
  // per-medium field index, include/compute/fields/impl/field_impl.tpp:39-62
  // (loop order chunk, ix, iz, element; elements of the other medium are
  // skipped).  Elastic elements are [0, nspec_el) in compute order.
  auto number_medium = [&](bool acoustic, std::vector<int32_t> &index) {
    index.assign(nglob, -1);
    int count = 0;
    for (int c0 = 0; c0 < nspec; c0 += chunk)
      for (int ix = 0; ix < NGLL; ix++)
        for (int iz = 0; iz < NGLL; iz++)
          for (int ie = 0; ie < chunk && c0 + ie < nspec; ie++) {
            if ((c0 + ie >= a.nspec_el) != acoustic) continue;
            const int g = a.index_mapping[static_cast<size_t>(c0 + ie) * npe + iz * NGLL + ix];
            if (index[g] < 0) index[g] = count++;
          }
    return count;
  };
  a.nglob_el = number_medium(false, a.field_index);
  a.nglob_ac = a.nspec_el < nspec ? number_medium(true, a.field_index_ac) : 0;
  if (a.nspec_el == nspec && a.nglob_el != nglob) throw std::runtime_error("medium numbering: count mismatch");
}

struct Locator2D {
  const HostAssembly &a;
  const std::vector<float> &ctrl;  // (nspec, ngnod, 2)
  const std::vector<float> &px, &pz;
  int ngnod;
  std::vector<float> box;              // (nspec, 4): x min, x max, z min, z max of the GLL points
  std::vector<uint32_t> corner_start;  // (nglob + 1) CSR over the elements having a global point as corner
  std::vector<int32_t> corner_elem;    // ascending element order within each point
  // uniform G x G grid over the points' bounding box; cell c lists (CSR,
  // ascending) every element whose bounding box overlaps it
  int G = 1;
  double X0 = 0, Z0 = 0, sx = 1, sz = 1;
  std::vector<uint32_t> cell_start;
  std::vector<int32_t> cell_elem;

  int cell_x(double x) const { return static_cast<int>(std::min(std::max(std::floor((x - X0) * sx), 0.0), G - 1.0)); }
  int cell_z(double z) const { return static_cast<int>(std::min(std::max(std::floor((z - Z0) * sz), 0.0), G - 1.0)); }

  Locator2D(const HostAssembly &a_, const std::vector<float> &ctrl_, const std::vector<float> &px_,
            const std::vector<float> &pz_, int ngnod_)
      : a(a_), ctrl(ctrl_), px(px_), pz(pz_), ngnod(ngnod_) {
    const int nspec = a.nspec;
    box.resize(4 * static_cast<size_t>(nspec));
    for (int e = 0; e < nspec; e++) {
      float x0 = px[static_cast<size_t>(e) * 25], x1 = x0, z0 = pz[static_cast<size_t>(e) * 25], z1 = z0;
      for (int p = 1; p < 25; p++) {
        const float x = px[static_cast<size_t>(e) * 25 + p], z = pz[static_cast<size_t>(e) * 25 + p];
        x0 = std::min(x0, x);
        x1 = std::max(x1, x);
        z0 = std::min(z0, z);
        z1 = std::max(z1, z);
      }
      float *b = &box[4 * static_cast<size_t>(e)];
      b[0] = x0;
      b[1] = x1;
      b[2] = z0;
      b[3] = z1;
    }
    double X1 = box[0], Z1 = box[2];
    X0 = box[0];
    Z0 = box[2];
    for (int e = 0; e < nspec; e++) {
      const float *b = &box[4 * static_cast<size_t>(e)];
      X0 = std::min<double>(X0, b[0]);
      X1 = std::max<double>(X1, b[1]);
      Z0 = std::min<double>(Z0, b[2]);
      Z1 = std::max<double>(Z1, b[3]);
    }
    G = std::max(1, static_cast<int>(std::sqrt(static_cast<double>(nspec))));
    sx = X1 > X0 ? G / (X1 - X0) : 0.0;
    sz = Z1 > Z0 ? G / (Z1 - Z0) : 0.0;
    cell_start.assign(static_cast<size_t>(G) * G + 1, 0);
    for (int pass = 0; pass < 2; pass++) {
      std::vector<uint32_t> fill;
      if (pass == 1) {
        for (size_t c = 0; c + 1 < cell_start.size(); c++) cell_start[c + 1] += cell_start[c];
        cell_elem.resize(cell_start.back());
        fill.assign(cell_start.begin(), cell_start.end() - 1);
      }
      for (int e = 0; e < nspec; e++) {
        const float *b = &box[4 * static_cast<size_t>(e)];
        for (int cz = cell_z(b[2]); cz <= cell_z(b[3]); cz++)
          for (int cx = cell_x(b[0]); cx <= cell_x(b[1]); cx++) {
            const size_t c = static_cast<size_t>(cz) * G + cx;
            if (pass == 0)
              cell_start[c + 1]++;
            else
              cell_elem[fill[c]++] = e;
          }
      }
    }
    corner_start.assign(static_cast<size_t>(a.nglob) + 1, 0);
    for (int e = 0; e < nspec; e++)
      for (int c : CORNERS) corner_start[a.index_mapping[static_cast<size_t>(e) * 25 + c] + 1]++;
    for (int g = 0; g < a.nglob; g++) corner_start[g + 1] += corner_start[g];
    corner_elem.resize(4 * static_cast<size_t>(nspec));
    std::vector<uint32_t> fill(corner_start.begin(), corner_start.end() - 1);
    for (int e = 0; e < nspec; e++)
      for (int c : CORNERS) corner_elem[fill[a.index_mapping[static_cast<size_t>(e) * 25 + c]]++] = e;
  }

  static constexpr int CORNERS[4] = {0, 4, 20, 24};

  void position(int e, float xi, float gamma, float &x, float &z) const {
    float s[9];
    shape2d(xi, gamma, ngnod, s);
    x = 0.0f;
    z = 0.0f;
    for (int n = 0; n < ngnod; n++) {
      x += s[n] * ctrl[(static_cast<size_t>(e) * ngnod + n) * 2];
      z += s[n] * ctrl[(static_cast<size_t>(e) * ngnod + n) * 2 + 1];
    }
  }
  void inverse_jacobian(int e, float xi, float gamma, float &xix, float &gammax, float &xiz, float &gammaz) const {
    float d0[9], d1[9];
    dshape2d(xi, gamma, ngnod, d0, d1);
    float xxi = 0.0f, zxi = 0.0f, xgamma = 0.0f, zgamma = 0.0f;
    for (int n = 0; n < ngnod; n++) {
      const float cx = ctrl[(static_cast<size_t>(e) * ngnod + n) * 2];
      const float cz = ctrl[(static_cast<size_t>(e) * ngnod + n) * 2 + 1];
      xxi += d0[n] * cx;
      zxi += d0[n] * cz;
      xgamma += d1[n] * cx;
      zgamma += d1[n] * cz;
    }
    const float J = xxi * zgamma - xgamma * zxi;
    xix = zgamma / J;
    gammax = -zxi / J;
    xiz = -xgamma / J;
    gammaz = xxi / J;
  }

  // src/algorithms/locate_point.cpp:120-176: nearest GLL point (strict <,
  // first in (element, iz, ix) order), candidate elements sharing a corner
  // point of the guess, 100 clamped Newton steps per candidate, smallest
  // distance wins.
  //
  // The nearest-point scan visits only elements that can hold a point within
  // reach of the target.  U, the squared distance (in double) to the centre
  // point of some element of the grid cells around the target, bounds the
  // minimum; reach = sqrt(U) (1 + 1e-5).  Every element whose bounding box
  // comes within reach overlaps a grid cell of the square of half-width
  // reach (1 + 1e-5) / (1 - 1e-5) about the target (cell indices are
  // monotonic in the coordinate), so it is in the candidate list, visited in
  // ascending element order.  A skipped element lies farther than reach, so
  // every float distance in it exceeds the float minimum (float distances
  // are within a few ulp of the exact ones), and the first minimum in scan
  // order is the one a scan over all elements finds.  The corner candidates
  // come from the corner index in ascending element order, as a scan over all
  // elements collects them.
  void locate(float gx, float gz, int &ispec, float &xi_out, float &gamma_out) const {
    if (!std::isfinite(gx) || !std::isfinite(gz)) throw std::invalid_argument("non-finite source or receiver position");
    const double X = gx, Z = gz;
    const int qx = cell_x(X), qz = cell_z(Z);
    double U = std::numeric_limits<double>::max();
    for (int r = 0; r <= G && U == std::numeric_limits<double>::max(); r++)
      for (int cz = std::max(qz - r, 0); cz <= std::min(qz + r, G - 1); cz++)
        for (int cx = std::max(qx - r, 0); cx <= std::min(qx + r, G - 1); cx++) {
          if (std::max(std::abs(cz - qz), std::abs(cx - qx)) != r) continue;
          const size_t c = static_cast<size_t>(cz) * G + cx;
          for (uint32_t k = cell_start[c]; k < cell_start[c + 1]; k++) {
            const size_t p = static_cast<size_t>(cell_elem[k]) * 25 + 12;
            const double dx = X - px[p], dz = Z - pz[p];
            U = std::min(U, dx * dx + dz * dz);
          }
        }
    const double reach = std::sqrt(U) * (1.0 + 1e-5), R = reach * (1.0 + 1e-5) / (1.0 - 1e-5);
    std::vector<int> near;
    for (int cz = cell_z(Z - R); cz <= cell_z(Z + R); cz++)
      for (int cx = cell_x(X - R); cx <= cell_x(X + R); cx++) {
        const size_t c = static_cast<size_t>(cz) * G + cx;
        near.insert(near.end(), cell_elem.begin() + cell_start[c], cell_elem.begin() + cell_start[c + 1]);
      }
    std::sort(near.begin(), near.end());
    near.erase(std::unique(near.begin(), near.end()), near.end());
    float dmin = std::numeric_limits<float>::max();
    int eg = 0, ixg = 0, izg = 0;
    for (int e : near) {
      const float *b = &box[4 * static_cast<size_t>(e)];
      const double bx = std::max({static_cast<double>(b[0]) - X, X - static_cast<double>(b[1]), 0.0});
      const double bz = std::max({static_cast<double>(b[2]) - Z, Z - static_cast<double>(b[3]), 0.0});
      if (std::sqrt(bx * bx + bz * bz) * (1.0 - 1e-5) > reach) continue;
      for (int j = 0; j < NGLL; j++)
        for (int i = 0; i < NGLL; i++) {
          const size_t p = static_cast<size_t>(e) * 25 + j * NGLL + i;
          const float dx = gx - px[p], dz = gz - pz[p];
          const float d = std::sqrt(dx * dx + dz * dz);
          if (d < dmin) {
            dmin = d;
            eg = e;
            ixg = i;
            izg = j;
          }
        }
    }
    std::vector<int> cand;
    for (int c : CORNERS) {
      const int g = a.index_mapping[static_cast<size_t>(eg) * 25 + c];
      for (uint32_t k = corner_start[g]; k < corner_start[g + 1]; k++)
        if (corner_elem[k] != eg) cand.push_back(corner_elem[k]);
    }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    cand.insert(cand.begin(), eg);
    float best = std::numeric_limits<float>::max();
    ispec = eg;
    xi_out = a.xi[ixg];
    gamma_out = a.xi[izg];
    for (int e : cand) {
      float xi = a.xi[ixg], gamma = a.xi[izg];
      for (int it = 0; it < 100; it++) {
        float x, z, xix, gammax, xiz, gammaz;
        position(e, xi, gamma, x, z);
        inverse_jacobian(e, xi, gamma, xix, gammax, xiz, gammaz);
        const float dx = -(x - gx), dz = -(z - gz);
        xi += xix * dx + xiz * dz;
        gamma += gammax * dx + gammaz * dz;
        xi = std::min(std::max(xi, -1.01f), 1.01f);
        gamma = std::min(std::max(gamma, -1.01f), 1.01f);
      }
      float x, z;
      position(e, xi, gamma, x, z);
      const float dx = gx - x, dz = gz - z;
      const float d = std::sqrt(dx * dx + dz * dz);
      if (d < best) {
        best = d;
        ispec = e;
        xi_out = xi;
        gamma_out = gamma;
      }
      ixg = NGLL / 2;
      izg = NGLL / 2;
    }
  }
};

// Stacey traction for a point (include/boundary_conditions/stacey/stacey.tpp:36-72)
// with the velocity replaced by a constant, as the mass-matrix term uses it
// (stacey.tpp:641-677).
template <int D>
void stacey_term(const float *n, float w, float rho_vp, float rho_vs, float vconst, float *out) {
  float vn = 0.0f, nn = 0.0f;
  for (int d = 0; d < D; d++) {
    vn += vconst * n[d];
    nn += n[d] * n[d];
  }
  const float j1 = std::sqrt(nn);
  for (int d = 0; d < D; d++) {
    const float f = ((vn * n[d] / (j1 * j1)) * (rho_vp - rho_vs)) + vconst * rho_vs;
    out[d] += -1.0f * f * j1 * w;
  }
}

void setup_time_function(HostAssembly &a, int nsrc, const sfx_source *src, const sfx_run_desc &run) {
  // Automatic start: t0 = min over sources of (-1.2 / f0 + tshift), and every
  // source's shift becomes its own t0 - min t0.  A user start time keeps the
  // sources' own tshift and must not be later than min t0 - min tshift
  // (src/io/sources.cpp:128-158, ricker.cpp:15-20).
  float min_t0 = std::numeric_limits<float>::max(), min_tshift = std::numeric_limits<float>::max();
  std::vector<float> t0s(nsrc);
  const float dtf = static_cast<float>(run.dt);
  // a Dirac's frequency is 1 / (10 dt) (src/source_time_function/dirac.cpp:24-48)
  auto f0_of = [&](int s) {
    return src[s].stf == SFX_STF_DIRAC ? static_cast<float>(1.0 / (10.0 * dtf)) : static_cast<float>(src[s].f0);
  };
  for (int s = 0; s < nsrc; s++) {
    if (src[s].stf != SFX_STF_RICKER && src[s].stf != SFX_STF_DIRAC)
      throw std::invalid_argument("source time function must be SFX_STF_RICKER or SFX_STF_DIRAC");
    if (src[s].stf == SFX_STF_RICKER && !(src[s].f0 > 0)) throw std::invalid_argument("source f0 must be positive");
    const float hdur = 1.0f / f0_of(s);
    t0s[s] = -1.2 * hdur + static_cast<float>(src[s].tshift);
    min_t0 = std::min(min_t0, t0s[s]);
    min_tshift = std::min(min_tshift, static_cast<float>(src[s].tshift));
  }
  if (nsrc == 0) min_t0 = min_tshift = 0.0f;
  const float user = static_cast<float>(run.start_time);
  const bool user_defined = std::abs(user) > std::numeric_limits<float>::epsilon();
  if (user_defined && user > min_t0 - min_tshift)
    throw std::invalid_argument("start_time is later than the earliest source onset allows");
  const float t0 = user_defined ? user : min_t0;
  a.t0 = t0;
  const float dt = static_cast<float>(run.dt);
  a.stf.assign(static_cast<size_t>(run.nstep) * nsrc, 0.0f);
  for (int s = 0; s < nsrc; s++) {
    const float shift = user_defined ? static_cast<float>(src[s].tshift) : t0s[s] - min_t0;
    for (int i = 0; i < run.nstep; i++)
      a.stf[static_cast<size_t>(i) * nsrc + s] =
          src[s].stf == SFX_STF_DIRAC
              ? dirac(t0 + i * dt - shift, f0_of(s), static_cast<float>(src[s].factor))
              : ricker(t0 + i * dt - shift, static_cast<float>(src[s].f0), static_cast<float>(src[s].factor));
  }
}

void check_run(const sfx_run_desc &run) {
  if (!(run.dt > 0)) throw std::invalid_argument("dt must be positive");
  if (run.nstep < 0) throw std::invalid_argument("nstep must be >= 0");
  if (run.nstep_between_samples < 1) throw std::invalid_argument("nstep_between_samples must be >= 1");
  if (run.nseismo_types < 0 || run.nseismo_types > 4) throw std::invalid_argument("nseismo_types must be 0..4");
  for (int i = 0; i < run.nseismo_types; i++)
    if (run.seismo_types[i] < 0 || run.seismo_types[i] > 3) throw std::invalid_argument("bad seismogram type");
  if (run.chunk_size < 1) throw std::invalid_argument("chunk_size must be >= 1");
  if (!std::isfinite(run.start_time)) throw std::invalid_argument("start_time must be finite");
  if (run.scatter != SFX_SCATTER_COLORED && run.scatter != SFX_SCATTER_ATOMIC && run.scatter != SFX_SCATTER_GATHER &&
      run.scatter != SFX_SCATTER_BLOCK)
    throw std::invalid_argument("scatter must be SFX_SCATTER_COLORED, _ATOMIC, _GATHER or _BLOCK");
}

}  // namespace

// ---------------------------------------------------------------------------
// 2-D
// ---------------------------------------------------------------------------
// Mass matrix and inverse (compute_mass_matrix.tpp:80-114,
// invert_mass_matrix.tpp:40-47), accumulated in element order, with the
// Stacey term at the points of the Stacey entries.
void mass_2d(HostAssembly &a, double dt) {
  const int npe = 25;
  std::vector<int> bnd_of(static_cast<size_t>(a.nspec) * npe, -1);
  for (size_t q = 0; q < a.bnd_slot.size(); q++) bnd_of[a.bnd_slot[q]] = static_cast<int>(q);
  const int ng = a.nglob_el;
  std::vector<float> mass(2 * static_cast<size_t>(ng), 0.0f);
  const float vfac = -1.0f * static_cast<float>(dt) * 0.5f;
  for (int e = 0; e < a.nspec_el; e++)
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        const size_t p = static_cast<size_t>(e) * npe + iz * NGLL + ix;
        float mm[2] = {a.rho[p], a.rho[p]};
        const float f = a.w[ix] * a.w[iz] * a.jac[p];
        mm[0] *= f;
        mm[1] *= f;
        const int k = bnd_of[p];
        if (k >= 0) stacey_term<2>(&a.bnd_normal[2 * k], a.bnd_weight[k], a.bnd_rho_vp[k], a.bnd_rho_vs[k], vfac, mm);
        const int gi = a.gather[p];
        mass[gi] += mm[0];
        mass[ng + gi] += mm[1];
      }
  a.mass_inverse.assign(2 * static_cast<size_t>(a.nglob_pad), 0.0f);
  for (int c = 0; c < 2; c++)
    for (int g = 0; g < ng; g++)
      a.mass_inverse[static_cast<size_t>(c) * a.nglob_pad + g] = 1.0f / mass[static_cast<size_t>(c) * ng + g];
  if (!a.coupled()) return;
  // acoustic: 1 / kappa (include/medium/dim2/acoustic/isotropic/
  // mass_matrix.tpp) times w_i w_j J, with the Stacey term of a constant
  // velocity -dt/2 (stacey.tpp:172-213 as the mass matrix uses it)
  std::vector<float> mass_ac(static_cast<size_t>(a.nglob_ac), 0.0f);
  for (int e = a.nspec_el; e < a.nspec; e++)
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        const size_t p = static_cast<size_t>(e) * npe + iz * NGLL + ix;
        const float f = a.w[ix] * a.w[iz] * a.jac[p];
        float m0 = 1.0f / a.kappa[p];
        m0 *= f;
        if (a.ac_stacey[p] != 0.0f) m0 += -1.0f * a.ac_stacey[p] * vfac;
        mass_ac[a.gather[p]] += m0;
      }
  a.mass_inverse_ac.assign(a.nglob_ac_pad, 0.0f);
  for (int g = 0; g < a.nglob_ac; g++) a.mass_inverse_ac[g] = 1.0f / mass_ac[g];
}

// New per-point properties (nspec, 25) in mesh element order: properties,
// Stacey impedances (stacey.cpp:64-181) and mass matrix are rebuilt.
void set_properties_2d(HostAssembly &a, const float *rho, const float *kappa, const float *mu, double dt) {
  if (a.dim != 2) throw std::invalid_argument("set_properties is implemented for 2-D meshes");
  if (a.coupled()) throw std::invalid_argument("set_properties is implemented for elastic meshes");
  if (!rho || !kappa || !mu) throw std::invalid_argument("null property array");
  const int npe = 25;
  for (int e = 0; e < a.nspec; e++)
    for (int p = 0; p < npe; p++) {
      const size_t src = static_cast<size_t>(a.compute_to_mesh[e]) * npe + p, dst = static_cast<size_t>(e) * npe + p;
      if (!(rho[src] > 0.0f) || !(mu[src] > 0.0f) || !(kappa[src] + (4.0f / 3.0f) * mu[src] > 0.0f))
        throw std::invalid_argument("properties must give rho > 0, mu > 0 and lambda + 2 mu > 0");
      a.rho[dst] = rho[src];
      a.kappa[dst] = kappa[src];
      a.mu[dst] = mu[src];
    }
  for (size_t q = 0; q < a.bnd_slot.size(); q++) {
    const size_t p = a.bnd_slot[q];
    const float lp2m = a.kappa[p] + static_cast<float>(4.0 / 3.0) * a.mu[p];
    a.bnd_rho_vp[q] = std::sqrt(lp2m / a.rho[p]) * a.rho[p];
    a.bnd_rho_vs[q] = std::sqrt(a.rho[p] * a.mu[p]);
  }
  mass_2d(a, dt);
}

namespace {

// Points of an edge in increasing order, (ix, iz)
// (include/compute/coupled_interfaces/interface_container.tpp:15-48).
void edge_point(int edge, int i, int &ix, int &iz) {
  switch (edge) {
    case SFX_EDGE_BOTTOM: ix = i; iz = 0; break;
    case SFX_EDGE_TOP: ix = i; iz = NGLL - 1; break;
    case SFX_EDGE_LEFT: ix = 0; iz = i; break;
    default: ix = NGLL - 1; iz = i; break;  // right
  }
}
// The coupling kernel's edge iterator (include/policies/edge.hpp, edge<dim2>):
// the point of its own element and of the coupled element at edge point ip.
void self_point(int edge, int ip, int &iz, int &ix) {
  const int n = NGLL - 1;
  switch (edge) {
    case SFX_EDGE_BOTTOM: iz = 0; ix = ip; break;
    case SFX_EDGE_TOP: iz = n; ix = n - ip; break;
    case SFX_EDGE_LEFT: iz = ip; ix = 0; break;
    default: iz = n - ip; ix = n; break;  // right
  }
}
void coupled_point(int edge, int ip, int &iz, int &ix) {
  const int n = NGLL - 1;
  switch (edge) {
    case SFX_EDGE_BOTTOM: iz = 0; ix = n - ip; break;
    case SFX_EDGE_TOP: iz = n; ix = ip; break;
    case SFX_EDGE_LEFT: iz = n - ip; ix = n; break;
    default: iz = ip; ix = 0; break;  // right
  }
}
// Edge normal scaled by the Jacobian (core/specfem/point/partial_derivatives.hpp:371-396).
void edge_normal(const HostAssembly &a, size_t p, int edge, float &nx, float &nz) {
  const float J = a.jac[p];
  switch (edge) {
    case SFX_EDGE_BOTTOM: nx = -1.0f * a.gammax[p] * J; nz = -1.0f * a.gammaz[p] * J; break;
    case SFX_EDGE_TOP: nx = a.gammax[p] * J; nz = a.gammaz[p] * J; break;
    case SFX_EDGE_LEFT: nx = -1.0f * a.xix[p] * J; nz = -1.0f * a.xiz[p] * J; break;
    default: nx = a.xix[p] * J; nz = a.xiz[p] * J; break;  // right
  }
}

}  // namespace

// Elastic-acoustic interface (include/compute/coupled_interfaces/
// interface_container.tpp:50-365): for every fluid-solid element pair the one
// pair of edges whose points coincide (distance over the elastic element's
// edge spacing <= 1e-10), the GLL weight along the edge as the integration
// factor and the elastic element's edge normals; then the entries of the two
// coupling terms in the order of the coupling kernel's edge iterator
// (coupled_interface.tpp): the acoustic side adds factor (n . u) with n the
// negated elastic normal, the elastic side factor n chi_dot_dot.
static void build_coupling_2d(HostAssembly &a, const sfx_mesh2d &m, const std::vector<int> &mesh_to_compute,
                              const std::vector<float> &px, const std::vector<float> &pz) {
  const int edges[4] = {SFX_EDGE_BOTTOM, SFX_EDGE_TOP, SFX_EDGE_LEFT, SFX_EDGE_RIGHT};
  auto pidx = [](int e, int iz, int ix) { return static_cast<size_t>(e) * 25 + iz * NGLL + ix; };
  for (int i = 0; i < m.nfs; i++) {
    if (m.fs_acoustic[i] < 0 || m.fs_acoustic[i] >= m.nspec || m.fs_elastic[i] < 0 || m.fs_elastic[i] >= m.nspec)
      throw std::invalid_argument("fluid-solid element out of range");
    const int e_el = mesh_to_compute[m.fs_elastic[i]], e_ac = mesh_to_compute[m.fs_acoustic[i]];
    if (!(e_el < a.nspec_el && e_ac >= a.nspec_el))
      throw std::invalid_argument("Coupled Interfaces: Interface is not between the correct mediums");
    int edge_el = -1, edge_ac = -1, found = 0;
    for (int ed1 : edges)
      for (int ed2 : edges) {
        if (ed1 == ed2) continue;
        const float edge_size = (ed1 == SFX_EDGE_BOTTOM || ed1 == SFX_EDGE_TOP)
                                    ? px[pidx(e_el, 0, 1)] - px[pidx(e_el, 0, 0)]
                                    : pz[pidx(e_el, 1, 0)] - pz[pidx(e_el, 0, 0)];
        bool connected = true;
        for (int ip = 0; ip < NGLL && connected; ip++) {
          int ix1, iz1, ix2, iz2;
          edge_point(ed1, ip, ix1, iz1);
          edge_point(ed2, ip, ix2, iz2);
          const size_t p1 = pidx(e_el, iz1, ix1), p2 = pidx(e_ac, iz2, ix2);
          const float dx = px[p1] - px[p2], dz = pz[p1] - pz[p2];
          if (std::sqrt(dx * dx + dz * dz) / edge_size > 1.e-10f) connected = false;
        }
        if (connected) {
          edge_el = ed1;
          edge_ac = ed2;
          found++;
        }
      }
    if (found != 1) throw std::invalid_argument("Number of connected edges is not 1");
    float factor[NGLL], nx[NGLL], nz[NGLL];
    for (int ip = 0; ip < NGLL; ip++) {
      int ix1, iz1, ix2, iz2;
      edge_point(edge_el, ip, ix1, iz1);
      edge_point(edge_ac, ip, ix2, iz2);
      float n1x, n1z, n2x, n2z;
      edge_normal(a, pidx(e_el, iz1, ix1), edge_el, n1x, n1z);
      edge_normal(a, pidx(e_ac, iz2, ix2), edge_ac, n2x, n2z);
      if ((std::abs(n1x + n2x) > 1.e-4f * (std::abs(n1x - n2x) / 2)) &&
          (std::abs(n1z + n2z) > 1.e-4f * (std::abs(n1z - n2z) / 2)))
        throw std::invalid_argument("Edge normals need to be opposite in direction");
      factor[ip] = (edge_el == SFX_EDGE_BOTTOM || edge_el == SFX_EDGE_TOP) ? a.w[ix1] : a.w[iz1];
      nx[ip] = n1x;
      nz[ip] = n1z;
    }
    for (int ip = 0; ip < NGLL; ip++) {
      int sz, sx, cz, cx;
      self_point(edge_ac, ip, sz, sx);
      coupled_point(edge_el, ip, cz, cx);
      a.cpl_ac_self.push_back(a.gather[pidx(e_ac, sz, sx)]);
      a.cpl_ac_other.push_back(a.gather[pidx(e_el, cz, cx)]);
      a.cpl_ac_coef.insert(a.cpl_ac_coef.end(), {factor[ip], -nx[ip], -nz[ip]});
      self_point(edge_el, ip, sz, sx);
      coupled_point(edge_ac, ip, cz, cx);
      a.cpl_el_self.push_back(a.gather[pidx(e_el, sz, sx)]);
      a.cpl_el_other.push_back(a.gather[pidx(e_ac, cz, cx)]);
      a.cpl_el_coef.insert(a.cpl_el_coef.end(), {factor[ip], nx[ip], nz[ip]});
    }
  }
}

// 2-D gather path: device points in row order of their coordinates (z, then
// x; z taken on a grid of 1e-6 of the mesh height, so the points of one
// lattice row of a structured mesh stay together).  The reference's point
// numbering follows the elements, so an element's rows of points lie far
// apart in it; in row order, the elements of one row of the mesh read and
// write a few contiguous runs of points.  gather, the Stacey point list and
// M^-1 move to device indices; lattice_of[f] is the device index of field
// point f (the API keeps the reference numbering).
static void order_points_by_rows(HostAssembly &a) {
  const int ng = a.nglob_el;
  const size_t nloc = static_cast<size_t>(a.nspec) * a.npe;
  std::vector<double> cx(ng, 0.0), cz(ng, 0.0);
  double zlo = 0.0, zhi = 0.0;
  for (size_t p = 0; p < nloc; p++) {
    cx[a.gather[p]] = a.px2d[p];
    cz[a.gather[p]] = a.pz2d[p];
    zlo = p ? std::min(zlo, static_cast<double>(a.pz2d[p])) : a.pz2d[p];
    zhi = p ? std::max(zhi, static_cast<double>(a.pz2d[p])) : a.pz2d[p];
  }
  const double grid = std::max(1e-6 * (zhi - zlo), 1e-30);
  std::vector<int64_t> row(ng);
  for (int g = 0; g < ng; g++) row[g] = std::llround((cz[g] - zlo) / grid);
  std::vector<int32_t> order(ng);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t u, int32_t v) {
    return row[u] != row[v] ? row[u] < row[v] : cx[u] < cx[v];
  });
  a.lattice_of.assign(ng, -1);
  for (int d = 0; d < ng; d++) a.lattice_of[order[d]] = d;
  for (auto &g : a.gather) g = a.lattice_of[g];
  for (auto &g : a.bnd_gather) g = a.lattice_of[g];
  std::vector<float> m(a.mass_inverse.size(), 0.0f);
  for (int c = 0; c < a.ncomp; c++)
    for (int g = 0; g < ng; g++)
      m[static_cast<size_t>(c) * a.nglob_pad + a.lattice_of[g]] = a.mass_inverse[static_cast<size_t>(c) * a.nglob_pad + g];
  a.mass_inverse = std::move(m);
}

HostAssembly build_assembly_2d(const sfx_mesh2d &m, int nsrc, const sfx_source *src, int nrec,
                               const sfx_receiver *rec, const sfx_run_desc &run) {
  check_run(run);
  if (run.scatter == SFX_SCATTER_BLOCK) throw std::invalid_argument("SFX_SCATTER_BLOCK is for 3-D meshes");
  if (m.nspec <= 0 || m.npgeo <= 0) throw std::invalid_argument("empty mesh");
  // element points are addressed as e * 25 + p in 32-bit device indices
  if (static_cast<int64_t>(m.nspec) * 25 > (int64_t(1) << 31) - 1)
    throw std::invalid_argument("2-D mesh too large: nspec x 25 points exceed 32-bit device indexing");
  if (m.ngnod != 4 && m.ngnod != 9) throw std::invalid_argument("ngnod must be 4 or 9");
  if (!m.coorg || !m.knods || !m.material || !m.rho || !m.vp || !m.vs) throw std::invalid_argument("null mesh array");
  // vs > 0: elastic P-SV isotropic; vs = 0: acoustic isotropic
  // (read_material_properties.cpp:172-238)
  for (int i = 0; i < m.nmaterials; i++)
    if (!(m.vs[i] >= 0) || !(m.rho[i] > 0) || !(m.vp[i] > 0))
      throw std::invalid_argument("materials need rho > 0, vp > 0 and vs >= 0 (vs = 0: acoustic)");
  if ((m.nafs && (!m.afs_ispec || !m.afs_type)) || (m.nfs && (!m.fs_acoustic || !m.fs_elastic)) || m.nafs < 0 ||
      m.nfs < 0)
    throw std::invalid_argument("null acoustic free-surface or interface array");
  std::vector<int> acoustic_mesh(m.nspec, 0);
  for (int e = 0; e < m.nspec; e++) {
    const int mat = m.material[e];
    if (mat < 0 || mat >= m.nmaterials) throw std::invalid_argument("material index out of range");
    acoustic_mesh[e] = m.vs[mat] == 0.0 ? 1 : 0;
  }
  HostAssembly a;
  a.dim = 2;
  a.ncomp = 2;
  a.npe = 25;
  a.nspec = m.nspec;
  const Gll g = make_gll();
  std::copy(g.xi, g.xi + 5, a.xi);
  std::copy(g.w, g.w + 5, a.w);
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) a.hprime[i][j] = g.hprime[i][j];

  std::vector<int> tag_mesh;
  compute_ordering(m, acoustic_mesh, a, tag_mesh);
  const int nspec = m.nspec, ngnod = m.ngnod, npe = 25;
  const bool coupled = a.nspec_el < nspec;
  if (coupled && (run.scatter == SFX_SCATTER_GATHER))
    throw std::invalid_argument("meshes with acoustic elements need SFX_SCATTER_COLORED or _ATOMIC");

  // control nodes in compute order, as type_real (compute_mesh.cpp:172-197)
  std::vector<float> ctrl(static_cast<size_t>(nspec) * ngnod * 2);
  for (int e = 0; e < nspec; e++) {
    const int em = a.compute_to_mesh[e];
    for (int n = 0; n < ngnod; n++) {
      const int node = m.knods[static_cast<size_t>(em) * ngnod + n];
      if (node < 0 || node >= m.npgeo) throw std::invalid_argument("knods out of range");
      ctrl[(static_cast<size_t>(e) * ngnod + n) * 2] = static_cast<float>(m.coorg[2 * static_cast<size_t>(node)]);
      ctrl[(static_cast<size_t>(e) * ngnod + n) * 2 + 1] =
          static_cast<float>(m.coorg[2 * static_cast<size_t>(node) + 1]);
    }
  }
  // GLL point coordinates: float products accumulated in double, stored as
  // float (compute_mesh.cpp:299-333).
  const size_t nloc = static_cast<size_t>(nspec) * npe;
  std::vector<float> px(nloc), pz(nloc);
  float shp[5][5][9];
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) shape2d(a.xi[ix], a.xi[iz], ngnod, shp[iz][ix]);
  for (int e = 0; e < nspec; e++)
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        double x = 0.0, z = 0.0;
        for (int n = 0; n < ngnod; n++) {
          x += ctrl[(static_cast<size_t>(e) * ngnod + n) * 2] * shp[iz][ix][n];
          z += ctrl[(static_cast<size_t>(e) * ngnod + n) * 2 + 1] * shp[iz][ix][n];
        }
        const size_t p = static_cast<size_t>(e) * npe + iz * NGLL + ix;
        px[p] = static_cast<float>(x);
        pz[p] = static_cast<float>(z);
      }
  number_points_2d(a, px, pz, run.chunk_size);
  // a field component is addressed as c * nglob_pad + point (c < 2)
  if (2 * ((static_cast<int64_t>(a.nglob) + 63) / 64 * 64) > (int64_t(1) << 31) - 1)
    throw std::invalid_argument("2-D mesh too large: 2 x nglob exceeds 32-bit device indexing");
  a.nglob_pad = (std::max(a.nglob_el, 1) + 63) / 64 * 64;
  a.nglob_ac_pad = coupled ? (std::max(a.nglob_ac, 1) + 63) / 64 * 64 : 0;
  // field index of every local point in its element's medium field
  a.gather.resize(nloc);
  for (size_t p = 0; p < nloc; p++)
    a.gather[p] = static_cast<int>(p / npe) < a.nspec_el ? a.field_index[a.index_mapping[p]]
                                                         : a.field_index_ac[a.index_mapping[p]];

  // Geometry (compute_partial_derivatives.cpp:46-91; jacobian.cpp:232-257)
  a.xix.resize(nloc);
  a.xiz.resize(nloc);
  a.gammax.resize(nloc);
  a.gammaz.resize(nloc);
  a.jac.resize(nloc);
  float dsh0[5][5][9], dsh1[5][5][9];
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) dshape2d(a.xi[ix], a.xi[iz], ngnod, dsh0[iz][ix], dsh1[iz][ix]);
  a.dsh2d.assign(static_cast<size_t>(npe) * 2 * ngnod, 0.0f);
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++)
      for (int n = 0; n < ngnod; n++) {
        a.dsh2d[(static_cast<size_t>(iz * NGLL + ix) * 2) * ngnod + n] = dsh0[iz][ix][n];
        a.dsh2d[(static_cast<size_t>(iz * NGLL + ix) * 2 + 1) * ngnod + n] = dsh1[iz][ix][n];
      }
  for (int e = 0; e < nspec; e++)
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        float xxi = 0.0f, zxi = 0.0f, xgamma = 0.0f, zgamma = 0.0f;
        for (int n = 0; n < ngnod; n++) {
          const float cx = ctrl[(static_cast<size_t>(e) * ngnod + n) * 2];
          const float cz = ctrl[(static_cast<size_t>(e) * ngnod + n) * 2 + 1];
          xxi += dsh0[iz][ix][n] * cx;
          zxi += dsh0[iz][ix][n] * cz;
          xgamma += dsh1[iz][ix][n] * cx;
          zgamma += dsh1[iz][ix][n] * cz;
        }
        const float J = xxi * zgamma - xgamma * zxi;
        if (J < 1e-10f)  // check_small_jacobian, compute_partial_derivatives.cpp:114-150
          throw std::runtime_error("element " + std::to_string(a.compute_to_mesh[e]) + " has a small jacobian");
        const size_t p = static_cast<size_t>(e) * npe + iz * NGLL + ix;
        a.xix[p] = zgamma / J;
        a.gammax[p] = -zxi / J;
        a.xiz[p] = -xgamma / J;
        a.gammaz[p] = xxi / J;
        a.jac[p] = J;
      }
  // constant geometry per element (bitwise) -> one record per element
  {
    bool regular = true;
    for (int e = 0; e < nspec && regular; e++) {
      const size_t b = static_cast<size_t>(e) * npe;
      for (int p = 1; p < npe && regular; p++)
        regular = a.xix[b + p] == a.xix[b] && a.xiz[b + p] == a.xiz[b] && a.gammax[b + p] == a.gammax[b] &&
                  a.gammaz[b + p] == a.gammaz[b] && a.jac[b + p] == a.jac[b];
    }
    a.geo2d.clear();
    if (regular) {
      a.geo2d.assign(static_cast<size_t>(nspec) * 8, 0.0f);
      for (int e = 0; e < nspec; e++) {
        const size_t b = static_cast<size_t>(e) * npe;
        float *g = &a.geo2d[static_cast<size_t>(e) * 8];
        g[0] = a.xix[b];
        g[1] = a.xiz[b];
        g[2] = a.gammax[b];
        g[3] = a.gammaz[b];
        g[4] = a.jac[b];
      }
    }
  }

  // Properties (include/medium/dim2/elastic/isotropic/material.hpp:62-77;
  // acoustic: kappa = rho cp^2, rho_inverse = 1 / rho,
  // include/medium/dim2/acoustic/isotropic/material.hpp:44-101)
  a.kappa.resize(nloc);
  a.mu.resize(nloc);
  a.rho.resize(nloc);
  if (coupled) a.rho_inv.assign(nloc, 0.0f);
  for (int e = 0; e < nspec; e++) {
    const int mat = m.material[a.compute_to_mesh[e]];
    const float rho = static_cast<float>(m.rho[mat]);
    const float cs = static_cast<float>(m.vs[mat]), cp = static_cast<float>(m.vp[mat]);
    const bool ac = e >= a.nspec_el;
    const float mu = ac ? 0.0f : rho * cs * cs;
    const float kappa = ac ? rho * cp * cp : static_cast<float>(rho * (cp * cp - (4.0 / 3.0) * cs * cs));
    for (int p = 0; p < npe; p++) {
      a.kappa[static_cast<size_t>(e) * npe + p] = kappa;
      a.mu[static_cast<size_t>(e) * npe + p] = mu;
      a.rho[static_cast<size_t>(e) * npe + p] = rho;
      if (ac) a.rho_inv[static_cast<size_t>(e) * npe + p] = 1.0f / rho;
    }
  }

  // Stacey entries (src/compute/boundaries/impl/stacey.cpp:64-181): for every
  // element point on an absorbing edge, the outward normal scaled by the
  // jacobian and the edge weight; the last listed edge of an element wins at
  // a point shared by two of its edges.
  std::vector<int> mesh_to_compute(nspec);
  for (int e = 0; e < nspec; e++) mesh_to_compute[a.compute_to_mesh[e]] = e;
  std::map<int, std::vector<int>> edges_of;
  for (int i = 0; i < m.nabs; i++) edges_of[mesh_to_compute[m.abs_ispec[i]]].push_back(i);
  std::vector<int> bnd_of(nloc, -1);  // element point -> Stacey entry
  if (coupled) a.ac_stacey.assign(nloc, 0.0f);
  for (const auto &kv : edges_of) {
    const int e = kv.first;
    if (e >= a.nspec_el) {
      // acoustic Stacey (stacey.tpp:172-213): w |n| rho_vpinverse per point,
      // rho_vpinverse = sqrt(rho_inverse / kappa) (core/specfem/point/
      // properties.hpp:68-76); normal and weight of the last listed edge
      float nxs[25], nzs[25], ws[25];
      bool on_any[25] = {};
      for (int i : kv.second) {
        const int edge = m.abs_edge[i];
        for (int iz = 0; iz < NGLL; iz++)
          for (int ix = 0; ix < NGLL; ix++) {
            const bool on = (edge == SFX_EDGE_TOP && iz == NGLL - 1) || (edge == SFX_EDGE_BOTTOM && iz == 0) ||
                            (edge == SFX_EDGE_LEFT && ix == 0) || (edge == SFX_EDGE_RIGHT && ix == NGLL - 1);
            if (!on) continue;
            const int lp = iz * NGLL + ix;
            const size_t p = static_cast<size_t>(e) * npe + lp;
            const float J = a.jac[p];
            if (edge == SFX_EDGE_LEFT || edge == SFX_EDGE_RIGHT) {
              const float sg = edge == SFX_EDGE_LEFT ? -1.0f : 1.0f;
              nxs[lp] = sg * a.xix[p] * J;
              nzs[lp] = sg * a.xiz[p] * J;
              ws[lp] = a.w[iz];
            } else {
              const float sg = edge == SFX_EDGE_BOTTOM ? -1.0f : 1.0f;
              nxs[lp] = sg * a.gammax[p] * J;
              nzs[lp] = sg * a.gammaz[p] * J;
              ws[lp] = a.w[ix];
            }
            on_any[lp] = true;
          }
      }
      for (int lp = 0; lp < 25; lp++) {
        if (!on_any[lp]) continue;
        const size_t p = static_cast<size_t>(e) * npe + lp;
        const float factor = ws[lp] * std::sqrt(nxs[lp] * nxs[lp] + nzs[lp] * nzs[lp]);
        const float kappa_inverse = 1.0f / a.kappa[p];
        const float rho_vpinverse = std::sqrt(a.rho_inv[p] * kappa_inverse);
        a.ac_stacey[p] = factor * rho_vpinverse;
      }
      continue;
    }
    int slot[25];
    std::fill(slot, slot + 25, -1);
    for (int i : kv.second) {
      const int edge = m.abs_edge[i];
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const bool on = (edge == SFX_EDGE_TOP && iz == NGLL - 1) || (edge == SFX_EDGE_BOTTOM && iz == 0) ||
                          (edge == SFX_EDGE_LEFT && ix == 0) || (edge == SFX_EDGE_RIGHT && ix == NGLL - 1);
          if (!on) continue;
          const int lp = iz * NGLL + ix;
          const size_t p = static_cast<size_t>(e) * npe + lp;
          const float J = a.jac[p];
          float nx, nz, w;
          if (edge == SFX_EDGE_LEFT) {
            nx = -1.0f * a.xix[p] * J;
            nz = -1.0f * a.xiz[p] * J;
            w = a.w[iz];
          } else if (edge == SFX_EDGE_RIGHT) {
            nx = a.xix[p] * J;
            nz = a.xiz[p] * J;
            w = a.w[iz];
          } else if (edge == SFX_EDGE_TOP) {
            nx = a.gammax[p] * J;
            nz = a.gammaz[p] * J;
            w = a.w[ix];
          } else {
            nx = -1.0f * a.gammax[p] * J;
            nz = -1.0f * a.gammaz[p] * J;
            w = a.w[ix];
          }
          int k = slot[lp];
          if (k < 0) {
            k = static_cast<int>(a.bnd_gather.size());
            slot[lp] = k;
            bnd_of[p] = k;
            a.bnd_gather.push_back(a.gather[p]);
            a.bnd_slot.push_back(static_cast<int32_t>(p));
            a.bnd_normal.push_back(0);
            a.bnd_normal.push_back(0);
            a.bnd_weight.push_back(0);
            const float kap = a.kappa[p], mu = a.mu[p], rho = a.rho[p];
            const float lp2m = kap + static_cast<float>(4.0 / 3.0) * mu;
            a.bnd_rho_vp.push_back(std::sqrt(lp2m / rho) * rho);
            a.bnd_rho_vs.push_back(std::sqrt(rho * mu));
          }
          a.bnd_normal[2 * k] = nx;
          a.bnd_normal[2 * k + 1] = nz;
          a.bnd_weight[k] = w;
        }
    }
  }

  if (coupled) {
    // acoustic free-surface points (src/compute/boundaries/impl/
    // acoustic_free_surface.cpp:10-103): the edge's points, or the corner
    a.ac_dirichlet.assign(nloc, 0);
    for (int i = 0; i < m.nafs; i++) {
      const int e = mesh_to_compute[m.afs_ispec[i]], type = m.afs_type[i], n = NGLL - 1;
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const bool on = (type == SFX_EDGE_TOP && iz == n) || (type == SFX_EDGE_BOTTOM && iz == 0) ||
                          (type == SFX_EDGE_LEFT && ix == 0) || (type == SFX_EDGE_RIGHT && ix == n) ||
                          (type == SFX_CORNER_BOTTOM_RIGHT && iz == 0 && ix == n) ||
                          (type == SFX_CORNER_BOTTOM_LEFT && iz == 0 && ix == 0) ||
                          (type == SFX_CORNER_TOP_RIGHT && iz == n && ix == n) ||
                          (type == SFX_CORNER_TOP_LEFT && iz == n && ix == 0);
          if (on) a.ac_dirichlet[static_cast<size_t>(e) * npe + iz * NGLL + ix] = 1;
        }
    }
    build_coupling_2d(a, m, mesh_to_compute, px, pz);
  } else if (m.nafs || m.nfs) {
    throw std::invalid_argument("free-surface or fluid-solid entries on a mesh without acoustic elements");
  }
  mass_2d(a, run.dt);

  // Sources and receivers (locate_point; force_source.cpp:70-78;
  // moment_tensor_source.cpp:88-130; compute_receivers.cpp:59-83)
  Locator2D loc{a, ctrl, px, pz, ngnod};
  setup_time_function(a, nsrc, src, run);
  a.nsrc = nsrc;
  a.src_ispec.resize(nsrc);
  a.src_array.assign(static_cast<size_t>(nsrc) * 2 * npe, 0.0f);
  a.src_local.resize(2 * static_cast<size_t>(nsrc));
  const float pi = static_cast<float>(M_PI);
  for (int s = 0; s < nsrc; s++) {
    int e;
    float xi, gamma;
    loc.locate(static_cast<float>(src[s].x), static_cast<float>(src[s].z), e, xi, gamma);
    a.src_ispec[s] = e;
    a.src_local[2 * s] = xi;
    a.src_local[2 * s + 1] = gamma;
    float hx[5], hpx[5], hg[5], hpg[5];
    lagrange<float>(xi, a.xi, hx, hpx);
    lagrange<float>(gamma, a.xi, hg, hpg);
    float *arr = &a.src_array[static_cast<size_t>(s) * 2 * npe];
    if (e >= a.nspec_el) {
      // in an acoustic element a force source is the Lagrange product alone
      // (force_source.cpp:51-56; the acoustic source term divides by kappa,
      // include/medium/dim2/acoustic/isotropic/source.hpp)
      if (src[s].kind != SFX_SOURCE_FORCE)
        throw std::invalid_argument("moment-tensor source in an acoustic element");
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) arr[iz * NGLL + ix] = hx[ix] * hg[iz];
    } else if (src[s].kind == SFX_SOURCE_FORCE) {
      const float ang = static_cast<float>(src[s].angle);
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const float hl = hx[ix] * hg[iz];
          arr[iz * NGLL + ix] = std::sin(pi / 180 * ang) * hl;
          arr[npe + iz * NGLL + ix] = -1.0 * std::cos(pi / 180 * ang) * hl;
        }
    } else if (src[s].kind == SFX_SOURCE_MOMENT_TENSOR) {
      float xixs = 0, gammaxs = 0, xizs = 0, gammazs = 0;
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const float hl = hx[ix] * hg[iz];
          const size_t p = static_cast<size_t>(e) * npe + iz * NGLL + ix;
          xixs += hl * a.xix[p];
          gammaxs += hl * a.gammax[p];
          xizs += hl * a.xiz[p];
          gammazs += hl * a.gammaz[p];
        }
      const float Mxx = static_cast<float>(src[s].m[0]), Mzz = static_cast<float>(src[s].m[1]),
                  Mxz = static_cast<float>(src[s].m[2]);
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const float dsx = (hpx[ix] * xixs) * hg[iz] + hx[ix] * (hpg[iz] * gammaxs);
          const float dsz = (hpx[ix] * xizs) * hg[iz] + hx[ix] * (hpg[iz] * gammazs);
          arr[iz * NGLL + ix] = Mxx * dsx + Mxz * dsz;
          arr[npe + iz * NGLL + ix] = Mxz * dsx + Mzz * dsz;
        }
    } else {
      throw std::invalid_argument("unknown source kind");
    }
  }
  // application order: element type (no boundary, then Stacey), then source
  // index (src/compute/compute_sources.cpp:232-270); on a mesh with acoustic
  // elements each medium's update applies its own sources, ordered by the
  // boundary tag of their element (none, free surface, Stacey, composite)
  for (int pass = 0; pass < 2; pass++)
    for (int s = 0; s < nsrc; s++)
      if ((a.src_ispec[s] >= a.nspec_none) == (pass == 1)) a.src_order.push_back(s);
  a.src_stf_kind.resize(nsrc);
  for (int s = 0; s < nsrc; s++) a.src_stf_kind[s] = src[s].stf;
  if (coupled) {
    std::vector<int> tag_c(nspec);
    for (int e = 0; e < nspec; e++) tag_c[e] = tag_mesh[a.compute_to_mesh[e]];
    for (int pass = 0; pass < 4; pass++)
      for (int s = 0; s < nsrc; s++)
        if (tag_c[a.src_ispec[s]] == pass) (a.src_ispec[s] < a.nspec_el ? a.src_order_el : a.src_order_ac).push_back(s);
  } else {
    a.src_order_el = a.src_order;
  }

  a.nrec = nrec;
  a.rec_ispec.resize(nrec);
  a.rec_h.resize(static_cast<size_t>(nrec) * npe);
  a.rec_local.resize(2 * static_cast<size_t>(nrec));
  for (int r = 0; r < nrec; r++) {
    int e;
    float xi, gamma;
    loc.locate(static_cast<float>(rec[r].x), static_cast<float>(rec[r].z), e, xi, gamma);
    a.rec_ispec[r] = e;
    a.rec_local[2 * r] = xi;
    a.rec_local[2 * r + 1] = gamma;
    float hx[5], hg[5];
    lagrange<float>(xi, a.xi, hx, nullptr);
    lagrange<float>(gamma, a.xi, hg, nullptr);
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) a.rec_h[static_cast<size_t>(r) * npe + iz * NGLL + ix] = hx[ix] * hg[iz];
    (e < a.nspec_el ? a.rec_el : a.rec_ac).push_back(r);
  }
  // kept for locating adjoint sources after the build
  a.ngnod2d = ngnod;
  a.ctrl2d = std::move(ctrl);
  a.px2d = std::move(px);
  a.pz2d = std::move(pz);
  if (run.scatter == SFX_SCATTER_GATHER && !coupled) order_points_by_rows(a);
  return a;
}

// Adjoint sources (src/source/adjoint_source.cpp:9-90): the Lagrange product
// at the located point for both components, each component driven by its own
// external time function (src/source_time_function/external.cpp).  Stored as
// two single-component sources per adjoint source (x, then z) so the source
// kernel's one time function per source applies.
void place_adjoint_sources_2d(HostAssembly &a, int nadj, const sfx_receiver *pos, const float *stf, int nstep) {
  if (a.dim != 2) throw std::invalid_argument("adjoint sources are implemented for 2-D meshes");
  if (a.coupled()) throw std::invalid_argument("adjoint runs are implemented for elastic meshes");
  if (nadj < 0 || (nadj && (!pos || !stf))) throw std::invalid_argument("null adjoint source array");
  const int npe = 25;
  Locator2D loc{a, a.ctrl2d, a.px2d, a.pz2d, a.ngnod2d};
  a.nadj = nadj;
  a.adj_ispec.assign(2 * static_cast<size_t>(nadj), 0);
  a.adj_array.assign(2 * static_cast<size_t>(nadj) * 2 * npe, 0.0f);
  a.adj_stf.assign(static_cast<size_t>(nstep) * 2 * nadj, 0.0f);
  for (int s = 0; s < nadj; s++) {
    int e;
    float xi, gamma;
    loc.locate(static_cast<float>(pos[s].x), static_cast<float>(pos[s].z), e, xi, gamma);
    float hx[5], hg[5];
    lagrange<float>(xi, a.xi, hx, nullptr);
    lagrange<float>(gamma, a.xi, hg, nullptr);
    for (int c = 0; c < 2; c++) {
      const int v = 2 * s + c;
      a.adj_ispec[v] = e;
      float *arr = &a.adj_array[static_cast<size_t>(v) * 2 * npe + static_cast<size_t>(c) * npe];
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) arr[iz * NGLL + ix] = hx[ix] * hg[iz];
      for (int i = 0; i < nstep; i++)
        a.adj_stf[static_cast<size_t>(i) * 2 * nadj + v] = stf[(static_cast<size_t>(s) * nstep + i) * 2 + c];
    }
  }
  a.adj_order.clear();
  for (int pass = 0; pass < 2; pass++)
    for (int v = 0; v < 2 * nadj; v++)
      if ((a.adj_ispec[v] >= a.nspec_none) == (pass == 1)) a.adj_order.push_back(v);
}

// ---------------------------------------------------------------------------
// 3-D structured hexahedra
// ---------------------------------------------------------------------------
namespace {

// Element containing a point, and its reference coordinates, on the
// structured grid: the exact inverse of x = x0 + (e + (1 + xi) / 2) h.  A point
// on an element face belongs to the element above it, except on the last face.
void locate_structured(double x, double x0, double h, int n, int &e, double &xi) {
  const double s = (x - x0) / h;
  int k = static_cast<int>(std::floor(s));
  k = std::min(std::max(k, 0), n - 1);
  e = k;
  xi = 2.0 * (s - k) - 1.0;
  xi = std::min(std::max(xi, -1.0), 1.0);
}

}  // namespace

void check_index_range_3d(int64_t nglob_pad, int64_t nspec) {
  const int64_t lim = (int64_t(1) << 31) - 1;
  if (3 * nglob_pad > lim)
    throw std::invalid_argument("3-D mesh too large: 3 x " + std::to_string(nglob_pad) +
                                " padded field points exceed 32-bit device indexing");
  if (nspec * 125 > lim)
    throw std::invalid_argument("3-D mesh too large: " + std::to_string(nspec) +
                                " elements x 125 points exceed 32-bit device indexing");
}

HostAssembly build_assembly_3d(const sfx_mesh3d &m, int nsrc, const sfx_source *src, int nrec,
                               const sfx_receiver *rec, const sfx_run_desc &run) {
  check_run(run);
  if (m.nx <= 0 || m.ny <= 0 || m.nz <= 0) throw std::invalid_argument("empty 3-D mesh");
  for (int i = 0; i < run.nseismo_types; i++)
    if (run.seismo_types[i] == SFX_PRESSURE) throw std::invalid_argument("pressure seismograms are 2-D only");
  if (run.scatter == SFX_SCATTER_BLOCK && ((m.nx | m.ny | m.nz) & 1))
    throw std::invalid_argument("SFX_SCATTER_BLOCK needs even element counts nx, ny, nz");
  if (!(m.hx > 0 && m.hy > 0 && m.hz > 0)) throw std::invalid_argument("element sizes must be positive");
  if (!m.material || !m.rho || !m.vp || !m.vs) throw std::invalid_argument("null mesh array");
  for (int i = 0; i < m.nmaterials; i++)
    if (!(m.vs[i] > 0) || !(m.rho[i] > 0) || !(m.vp[i] > 0))
      throw std::invalid_argument("only elastic isotropic materials (vs > 0) are supported");
  HostAssembly a;
  a.dim = 3;
  a.ncomp = 3;
  a.npe = 125;
  a.nx = m.nx;
  a.ny = m.ny;
  a.nz = m.nz;
  for (int f = 0; f < 6; f++)
    if (m.absorbing[f]) a.absorb_mask |= 1 << f;
  const int64_t nspec64 = static_cast<int64_t>(m.nx) * m.ny * m.nz;
  // Kernels address a field component as c * nglob_pad + point and an element
  // point as e * 125 + p in 32-bit integers: both must stay below 2^31.  The
  // check runs before any host array of the mesh's size is allocated.
  const bool block_layout = run.scatter == SFX_SCATTER_BLOCK;
  const int64_t TX = m.nx + 1, TY = m.ny + 1, TZ = m.nz + 1;
  const int nbx = m.nx / 2, nby = m.ny / 2, nbz = m.nz / 2;
  const int64_t nint64 = (static_cast<int64_t>(nbx) * nby * nbz * 343 + 63) / 64 * 64;
  const int64_t pad64 =
      block_layout ? (nint64 + static_cast<int64_t>(nbx + 1) * (nby + 1) * (nbz + 1) * OWNER_STRIDE + 63) / 64 * 64
                   : TX * TY * TZ * 64;
  check_index_range_3d(pad64, nspec64);
  a.nspec = static_cast<int>(nspec64);
  a.nspec_none = a.nspec;
  const Gll g = make_gll();
  std::copy(g.xi, g.xi + 5, a.xi);
  std::copy(g.w, g.w + 5, a.w);
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) a.hprime[i][j] = g.hprime[i][j];
  const int nspec = a.nspec, npe = 125;
  a.compute_to_mesh.resize(nspec);
  for (int e = 0; e < nspec; e++) a.compute_to_mesh[e] = e;

  // Numbering: first visit in (element, k, j, i) order over the lattice of GLL
  // points, the 3-D analogue of compute_mesh.cpp:45-170 with chunk size 1.
  // Points shared by neighbouring elements are identified by their lattice
  // position (4 ex + i, 4 ey + j, 4 ez + k), which is exact on this mesh.
  const int64_t LX = 4 * static_cast<int64_t>(m.nx) + 1, LY = 4 * static_cast<int64_t>(m.ny) + 1,
                LZ = 4 * static_cast<int64_t>(m.nz) + 1;
  const int64_t nlat = LX * LY * LZ;
  std::vector<int32_t> first(nlat, -1);
  const size_t nloc = static_cast<size_t>(nspec) * npe;
  a.index_mapping.resize(nloc);
  int inum = 0;
  for (int ez = 0; ez < m.nz; ez++)
    for (int ey = 0; ey < m.ny; ey++)
      for (int ex = 0; ex < m.nx; ex++) {
        const size_t e = (static_cast<size_t>(ez) * m.ny + ey) * m.nx + ex;
        for (int k = 0; k < 5; k++)
          for (int j = 0; j < 5; j++)
            for (int i = 0; i < 5; i++) {
              const int64_t L = ((4 * static_cast<int64_t>(ez) + k) * LY + (4 * ey + j)) * LX + (4 * ex + i);
              int32_t &f = first[L];
              if (f < 0) f = inum++;
              a.index_mapping[e * npe + (k * 5 + j) * 5 + i] = f;
            }
      }
  a.nglob = inum;
  a.nglob_el = inum;
  a.nspec_el = a.nspec;
  if (static_cast<int64_t>(inum) != nlat) throw std::runtime_error("3-D numbering: lattice not fully covered");
  a.field_index.resize(a.nglob);
  for (int i = 0; i < a.nglob; i++) a.field_index[i] = i;
  // Device fields are stored in 4x4x4 tiles of lattice points: tile (tx, ty, tz)
  // holds the points 4 tx <= I < 4 tx + 4 etc., x fastest inside the tile and
  // tiles x fastest.  One extra tile per axis holds the last lattice plane.  An
  // element's 64 points with local indices < 4 are one contiguous tile.
  // gather holds the device index of every local point; lattice_of maps the
  // reference numbering to it.
  // The block path (SFX_SCATTER_BLOCK) stores the fields in the block layout of
  // lattice.hpp instead: block interiors contiguous, then the surface points
  // grouped by owner block.
  if (block_layout) {
    a.nint_pad = static_cast<int>(nint64);
    a.nglob_pad = static_cast<int>(pad64);
  } else {
    a.nglob_pad = static_cast<int>(pad64);
  }
  auto tiled = [&](int64_t I, int64_t J, int64_t K) {
    if (block_layout)
      return static_cast<int32_t>(block_layout_index(nbx, nby, a.nint_pad, static_cast<int>(I), static_cast<int>(J),
                                                     static_cast<int>(K)));
    return static_cast<int32_t>((((K >> 2) * TY + (J >> 2)) * TX + (I >> 2)) * 64 + ((K & 3) * 4 + (J & 3)) * 4 + (I & 3));
  };
  a.lattice_of.assign(a.nglob, -1);
  a.gather.resize(nloc);
  for (int ez = 0; ez < m.nz; ez++)
    for (int ey = 0; ey < m.ny; ey++)
      for (int ex = 0; ex < m.nx; ex++) {
        const size_t e = (static_cast<size_t>(ez) * m.ny + ey) * m.nx + ex;
        for (int k = 0; k < 5; k++)
          for (int j = 0; j < 5; j++)
            for (int i = 0; i < 5; i++) {
              const int32_t d = tiled(4 * ex + i, 4 * ey + j, 4 * ez + k);
              const size_t p = e * npe + (k * 5 + j) * 5 + i;
              a.gather[p] = d;
              a.lattice_of[a.index_mapping[p]] = d;
            }
      }

  // Regular geometry: d(xi)/dx = 2 / hx etc., jacobian hx hy hz / 8
  // (the regular-element compression of the 3-D database,
  // src/io/mesh/impl/fortran/dim3/mesh.cpp:150-151).
  a.xix3 = static_cast<float>(2.0 / m.hx);
  a.ety3 = static_cast<float>(2.0 / m.hy);
  a.gaz3 = static_cast<float>(2.0 / m.hz);
  a.jac3 = static_cast<float>(m.hx * m.hy * m.hz / 8.0);

  // Properties per point (per-point kappa/mu/rho of the 3-D database,
  // mesh.cpp:208-209,265), kappa = rho (vp^2 - 4/3 vs^2), mu = rho vs^2.
  a.kappa.resize(nloc);
  a.mu.resize(nloc);
  a.rho.resize(nloc);
  for (int e = 0; e < nspec; e++) {
    const int mat = m.material[e];
    if (mat < 0 || mat >= m.nmaterials) throw std::invalid_argument("material index out of range");
    const float rho = static_cast<float>(m.rho[mat]);
    const float cs = static_cast<float>(m.vs[mat]), cp = static_cast<float>(m.vp[mat]);
    const float mu = rho * cs * cs;
    const float kappa = rho * (cp * cp - (4.0 / 3.0) * cs * cs);
    std::fill(&a.kappa[static_cast<size_t>(e) * npe], &a.kappa[static_cast<size_t>(e) * npe] + npe, kappa);
    std::fill(&a.mu[static_cast<size_t>(e) * npe], &a.mu[static_cast<size_t>(e) * npe] + npe, mu);
    std::fill(&a.rho[static_cast<size_t>(e) * npe], &a.rho[static_cast<size_t>(e) * npe] + npe, rho);
  }

  // Stacey faces: one entry per element point of every absorbing face (a point
  // on two absorbing faces of one element gets both terms, as the face lists
  // of the 3-D database do).  Normal = outward unit normal times the face
  // jacobian; weight = product of the two in-face GLL weights.
  const double fx = m.hy * m.hz / 4.0, fy = m.hx * m.hz / 4.0, fz = m.hx * m.hy / 4.0;
  auto add_face = [&](int e, int face) {
    for (int b = 0; b < 5; b++)
      for (int c = 0; c < 5; c++) {
        int i, j, k;
        float n[3] = {0, 0, 0};
        float w;
        switch (face) {
          case 0: i = 0; j = c; k = b; n[0] = static_cast<float>(-fx); w = a.w[j] * a.w[k]; break;
          case 1: i = 4; j = c; k = b; n[0] = static_cast<float>(fx); w = a.w[j] * a.w[k]; break;
          case 2: j = 0; i = c; k = b; n[1] = static_cast<float>(-fy); w = a.w[i] * a.w[k]; break;
          case 3: j = 4; i = c; k = b; n[1] = static_cast<float>(fy); w = a.w[i] * a.w[k]; break;
          case 4: k = 0; i = c; j = b; n[2] = static_cast<float>(-fz); w = a.w[i] * a.w[j]; break;
          default: k = 4; i = c; j = b; n[2] = static_cast<float>(fz); w = a.w[i] * a.w[j]; break;
        }
        const size_t p = static_cast<size_t>(e) * npe + (k * 5 + j) * 5 + i;
        a.bnd_gather.push_back(a.gather[p]);
        a.bnd_slot.push_back(static_cast<int32_t>(p));
        a.bnd_normal.push_back(n[0]);
        a.bnd_normal.push_back(n[1]);
        a.bnd_normal.push_back(n[2]);
        a.bnd_weight.push_back(w);
        const float kap = a.kappa[p], mu = a.mu[p], rho = a.rho[p];
        const float lp2m = kap + static_cast<float>(4.0 / 3.0) * mu;
        a.bnd_rho_vp.push_back(std::sqrt(lp2m / rho) * rho);
        a.bnd_rho_vs.push_back(std::sqrt(rho * mu));
      }
  };
  for (int ez = 0; ez < m.nz; ez++)
    for (int ey = 0; ey < m.ny; ey++)
      for (int ex = 0; ex < m.nx; ex++) {
        const int e = (ez * m.ny + ey) * m.nx + ex;
        if (m.absorbing[0] && ex == 0) add_face(e, 0);
        if (m.absorbing[1] && ex == m.nx - 1) add_face(e, 1);
        if (m.absorbing[2] && ey == 0) add_face(e, 2);
        if (m.absorbing[3] && ey == m.ny - 1) add_face(e, 3);
        if (m.absorbing[4] && ez == 0) add_face(e, 4);
        if (m.absorbing[5] && ez == m.nz - 1) add_face(e, 5);
      }

  // Group the entries of one element point (a point on two or three
  // absorbing faces of its element) so one thread adds them in a fixed order.
  {
    const size_t nb = a.bnd_slot.size();
    std::vector<size_t> order(nb);
    for (size_t q = 0; q < nb; q++) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return a.bnd_slot[x] < a.bnd_slot[y]; });
    auto reorder_i = [&](std::vector<int32_t> &v, int w) {
      std::vector<int32_t> o(v.size());
      for (size_t q = 0; q < nb; q++)
        for (int d = 0; d < w; d++) o[q * w + d] = v[order[q] * w + d];
      v.swap(o);
    };
    auto reorder_f = [&](std::vector<float> &v, int w) {
      std::vector<float> o(v.size());
      for (size_t q = 0; q < nb; q++)
        for (int d = 0; d < w; d++) o[q * w + d] = v[order[q] * w + d];
      v.swap(o);
    };
    reorder_i(a.bnd_gather, 1);
    reorder_i(a.bnd_slot, 1);
    reorder_f(a.bnd_normal, 3);
    reorder_f(a.bnd_weight, 1);
    reorder_f(a.bnd_rho_vp, 1);
    reorder_f(a.bnd_rho_vs, 1);
    a.bnd_seg.clear();
    for (size_t q = 0; q < nb; q++)
      if (q == 0 || a.bnd_slot[q] != a.bnd_slot[q - 1]) a.bnd_seg.push_back(static_cast<int32_t>(q));
    a.bnd_seg.push_back(static_cast<int32_t>(nb));
  }

  // Mass matrix with the implicit Stacey term (the 3-D form of
  // compute_mass_matrix.tpp:80-114 and stacey.tpp:641-677)
  {
    std::vector<float> mass(3 * static_cast<size_t>(a.nglob_pad), 0.0f);
    for (int e = 0; e < nspec; e++)
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++)
          for (int i = 0; i < 5; i++) {
            const size_t p = static_cast<size_t>(e) * npe + (k * 5 + j) * 5 + i;
            const float f = a.w[i] * a.w[j] * a.w[k] * a.jac3;
            const float mm = a.rho[p] * f;
            const int gi = a.gather[p];
            for (int c = 0; c < 3; c++) mass[static_cast<size_t>(c) * a.nglob_pad + gi] += mm;
          }
    const float vfac = -1.0f * static_cast<float>(run.dt) * 0.5f;
    for (size_t q = 0; q < a.bnd_gather.size(); q++) {
      float t[3] = {0, 0, 0};
      stacey_term<3>(&a.bnd_normal[3 * q], a.bnd_weight[q], a.bnd_rho_vp[q], a.bnd_rho_vs[q], vfac, t);
      for (int c = 0; c < 3; c++) mass[static_cast<size_t>(c) * a.nglob_pad + a.bnd_gather[q]] += t[c];
    }
    // padding slots (no lattice point) keep a zero inverse mass
    a.mass_inverse.assign(mass.size(), 0.0f);
    for (size_t i = 0; i < mass.size(); i++)
      if (mass[i] != 0.0f) a.mass_inverse[i] = 1.0f / mass[i];
  }

  // Sources (3-D force and moment tensor on the structured inverse map)
  setup_time_function(a, nsrc, src, run);
  a.nsrc = nsrc;
  a.src_ispec.resize(nsrc);
  a.src_array.assign(static_cast<size_t>(nsrc) * 3 * npe, 0.0f);
  a.src_local.resize(3 * static_cast<size_t>(nsrc));
  for (int s = 0; s < nsrc; s++) {
    int ex, ey, ez;
    double xi, eta, ga;
    locate_structured(src[s].x, m.x0, m.hx, m.nx, ex, xi);
    locate_structured(src[s].y, m.y0, m.hy, m.ny, ey, eta);
    locate_structured(src[s].z, m.z0, m.hz, m.nz, ez, ga);
    const int e = (ez * m.ny + ey) * m.nx + ex;
    a.src_ispec[s] = e;
    a.src_local[3 * s] = xi;
    a.src_local[3 * s + 1] = eta;
    a.src_local[3 * s + 2] = ga;
    float hx[5], hpx[5], hy[5], hpy[5], hz[5], hpz[5];
    lagrange<float>(static_cast<float>(xi), a.xi, hx, hpx);
    lagrange<float>(static_cast<float>(eta), a.xi, hy, hpy);
    lagrange<float>(static_cast<float>(ga), a.xi, hz, hpz);
    float *arr = &a.src_array[static_cast<size_t>(s) * 3 * npe];
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++)
        for (int i = 0; i < 5; i++) {
          const int p = (k * 5 + j) * 5 + i;
          if (src[s].kind == SFX_SOURCE_FORCE) {
            const float h = hx[i] * hy[j] * hz[k];
            arr[p] = static_cast<float>(src[s].fx) * h;
            arr[npe + p] = static_cast<float>(src[s].fy) * h;
            arr[2 * npe + p] = static_cast<float>(src[s].fz) * h;
          } else if (src[s].kind == SFX_SOURCE_MOMENT_TENSOR) {
            const float dx = hpx[i] * hy[j] * hz[k] * a.xix3;
            const float dy = hx[i] * hpy[j] * hz[k] * a.ety3;
            const float dz = hx[i] * hy[j] * hpz[k] * a.gaz3;
            const float Mxx = static_cast<float>(src[s].m[0]), Myy = static_cast<float>(src[s].m[1]),
                        Mzz = static_cast<float>(src[s].m[2]), Mxy = static_cast<float>(src[s].m[3]),
                        Mxz = static_cast<float>(src[s].m[4]), Myz = static_cast<float>(src[s].m[5]);
            arr[p] = Mxx * dx + Mxy * dy + Mxz * dz;
            arr[npe + p] = Mxy * dx + Myy * dy + Myz * dz;
            arr[2 * npe + p] = Mxz * dx + Myz * dy + Mzz * dz;
          } else {
            throw std::invalid_argument("unknown source kind");
          }
        }
  }
  for (int s = 0; s < nsrc; s++) a.src_order.push_back(s);

  a.nrec = nrec;
  a.rec_ispec.resize(nrec);
  a.rec_h.resize(static_cast<size_t>(nrec) * npe);
  a.rec_local.resize(3 * static_cast<size_t>(nrec));
  for (int r = 0; r < nrec; r++) {
    int ex, ey, ez;
    double xi, eta, ga;
    locate_structured(rec[r].x, m.x0, m.hx, m.nx, ex, xi);
    locate_structured(rec[r].y, m.y0, m.hy, m.ny, ey, eta);
    locate_structured(rec[r].z, m.z0, m.hz, m.nz, ez, ga);
    a.rec_ispec[r] = (ez * m.ny + ey) * m.nx + ex;
    a.rec_local[3 * r] = xi;
    a.rec_local[3 * r + 1] = eta;
    a.rec_local[3 * r + 2] = ga;
    float hx[5], hy[5], hz[5];
    lagrange<float>(static_cast<float>(xi), a.xi, hx, nullptr);
    lagrange<float>(static_cast<float>(eta), a.xi, hy, nullptr);
    lagrange<float>(static_cast<float>(ga), a.xi, hz, nullptr);
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++)
        for (int i = 0; i < 5; i++) a.rec_h[static_cast<size_t>(r) * npe + (k * 5 + j) * 5 + i] = hx[i] * hy[j] * hz[k];
  }
  return a;
}

// ---------------------------------------------------------------------------
// Element coloring
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// 3-D mesh database (general hexahedra)
// ---------------------------------------------------------------------------
namespace {

// Element and reference coordinates of a point on a 3-D database mesh: the
// nearest GLL point (squared distance in double, first minimum in global
// index order), then, in every element holding that point (ascending element
// order), Newton iterations on the isoparametric map of the element's 125 GLL
// points, x(xi) = sum_p l_p(xi) x_p, from the reference coordinates of that
// point, each step clamped to [-1.01, 1.01]; the candidate with the smallest
// final distance wins (first on ties).  The same steps as the 2-D
// locate_point (src/algorithms/locate_point.cpp:120-176) with the element's
// GLL points as its control nodes.
struct Locator3D {
  const sfx_mesh3d_db &m;
  std::vector<std::vector<int32_t>> owners;  // global point -> elements holding it (ascending)
  double xi_nodes[5];
  double tie = 0;  // final distances within this of the smallest are ties
  explicit Locator3D(const sfx_mesh3d_db &mesh, const float *nodes) : m(mesh), owners(mesh.nglob) {
    for (int e = 0; e < m.nspec; e++)
      for (int p = 0; p < 125; p++) {
        auto &o = owners[m.ibool[static_cast<size_t>(e) * 125 + p]];
        if (o.empty() || o.back() != e) o.push_back(e);
      }
    for (int i = 0; i < 5; i++) xi_nodes[i] = nodes[i];
    double lo[3] = {m.x[0], m.y[0], m.z[0]}, hi[3] = {m.x[0], m.y[0], m.z[0]};
    for (int g = 0; g < m.nglob; g++) {
      const double P[3] = {m.x[g], m.y[g], m.z[g]};
      for (int c = 0; c < 3; c++) {
        lo[c] = std::min(lo[c], P[c]);
        hi[c] = std::max(hi[c], P[c]);
      }
    }
    tie = 1e-9 * std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                           (hi[2] - lo[2]) * (hi[2] - lo[2]));
  }
  void locate(double x, double y, double z, int &ispec, double ref[3]) const {
    if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z))
      throw std::invalid_argument("non-finite source or receiver position");
    int best = -1;
    double dmin = std::numeric_limits<double>::max();
    for (int g = 0; g < m.nglob; g++) {
      const double dx = m.x[g] - x, dy = m.y[g] - y, dz = m.z[g] - z;
      const double d = dx * dx + dy * dy + dz * dz;
      if (d < dmin) {
        dmin = d;
        best = g;
      }
    }
    if (best < 0 || owners[best].empty()) throw std::runtime_error("locate: no element holds the nearest point");
    double final_best = std::numeric_limits<double>::max();
    ispec = -1;
    struct Cand {
      int e;
      double d, r[3];
    };
    std::vector<Cand> cand;
    for (int e : owners[best]) {
      // start from the reference coordinates of the nearest point in e
      double r[3] = {0, 0, 0};
      for (int p = 0; p < 125; p++)
        if (m.ibool[static_cast<size_t>(e) * 125 + p] == best) {
          r[0] = xi_nodes[p % 5];
          r[1] = xi_nodes[(p / 5) % 5];
          r[2] = xi_nodes[p / 25];
          break;
        }
      double d2 = 0;
      for (int it = 0; it < 100; it++) {
        double h[3][5], hp[3][5];
        for (int a = 0; a < 3; a++) lagrange<double>(r[a], xi_nodes, h[a], hp[a]);
        double X[3] = {0, 0, 0}, D[3][3] = {{0}};  // D[c][s] = d x_c / d r_s
        for (int k = 0; k < 5; k++)
          for (int j = 0; j < 5; j++)
            for (int i = 0; i < 5; i++) {
              const int g = m.ibool[static_cast<size_t>(e) * 125 + (k * 5 + j) * 5 + i];
              const double P[3] = {m.x[g], m.y[g], m.z[g]};
              const double w = h[0][i] * h[1][j] * h[2][k];
              const double dw[3] = {hp[0][i] * h[1][j] * h[2][k], h[0][i] * hp[1][j] * h[2][k],
                                    h[0][i] * h[1][j] * hp[2][k]};
              for (int c = 0; c < 3; c++) {
                X[c] += w * P[c];
                for (int s = 0; s < 3; s++) D[c][s] += dw[s] * P[c];
              }
            }
        const double res[3] = {x - X[0], y - X[1], z - X[2]};
        d2 = res[0] * res[0] + res[1] * res[1] + res[2] * res[2];
        const double det = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) -
                           D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                           D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
        if (det == 0.0) break;
        const double inv[3][3] = {
            {(D[1][1] * D[2][2] - D[1][2] * D[2][1]) / det, (D[0][2] * D[2][1] - D[0][1] * D[2][2]) / det,
             (D[0][1] * D[1][2] - D[0][2] * D[1][1]) / det},
            {(D[1][2] * D[2][0] - D[1][0] * D[2][2]) / det, (D[0][0] * D[2][2] - D[0][2] * D[2][0]) / det,
             (D[0][2] * D[1][0] - D[0][0] * D[1][2]) / det},
            {(D[1][0] * D[2][1] - D[1][1] * D[2][0]) / det, (D[0][1] * D[2][0] - D[0][0] * D[2][1]) / det,
             (D[0][0] * D[1][1] - D[0][1] * D[1][0]) / det}};
        for (int s = 0; s < 3; s++) {
          const double dr = inv[s][0] * res[0] + inv[s][1] * res[1] + inv[s][2] * res[2];
          r[s] = std::min(1.01, std::max(-1.01, r[s] + dr));
        }
      }
      // distance at the final reference coordinates
      {
        double h[3][5];
        for (int a = 0; a < 3; a++) lagrange<double>(r[a], xi_nodes, h[a], nullptr);
        double X[3] = {0, 0, 0};
        for (int k = 0; k < 5; k++)
          for (int j = 0; j < 5; j++)
            for (int i = 0; i < 5; i++) {
              const int g = m.ibool[static_cast<size_t>(e) * 125 + (k * 5 + j) * 5 + i];
              const double w = h[0][i] * h[1][j] * h[2][k];
              X[0] += w * m.x[g];
              X[1] += w * m.y[g];
              X[2] += w * m.z[g];
            }
        d2 = (x - X[0]) * (x - X[0]) + (y - X[1]) * (y - X[1]) + (z - X[2]) * (z - X[2]);
      }
      cand.push_back({e, std::sqrt(d2), {r[0], r[1], r[2]}});
      final_best = std::min(final_best, std::sqrt(d2));
    }
    // A point on a face, edge or corner shared by several elements is at the
    // same distance from each of them up to rounding: among the elements
    // within `tie` of the closest, the highest-numbered one holds it (on a
    // structured lattice, the element above the shared face in every
    // direction, as the structured mesh's location rule has it).
    for (const auto &c : cand)
      if (c.d <= final_best + tie && c.e > ispec) {
        ispec = c.e;
        for (int s = 0; s < 3; s++) ref[s] = c.r[s];
      }
  }
};

}  // namespace

HostAssembly build_assembly_3d_db(const sfx_mesh3d_db &m, int nsrc, const sfx_source *src, int nrec,
                                  const sfx_receiver *rec, const sfx_run_desc &run) {
  check_run(run);
  for (int i = 0; i < run.nseismo_types; i++)
    if (run.seismo_types[i] == SFX_PRESSURE) throw std::invalid_argument("pressure seismograms are 2-D only");
  if (run.scatter != SFX_SCATTER_COLORED && run.scatter != SFX_SCATTER_ATOMIC)
    throw std::invalid_argument("3-D database meshes need SFX_SCATTER_COLORED or _ATOMIC");
  if (m.nspec <= 0 || m.nglob <= 0) throw std::invalid_argument("empty 3-D mesh");
  if (!m.ibool || !m.x || !m.y || !m.z || !m.kappa || !m.mu || !m.rho || !m.rho_vp || !m.rho_vs || !m.ispec_type ||
      !m.jacobian || !m.xix || !m.xiy || !m.xiz || !m.etax || !m.etay || !m.etaz || !m.gammax || !m.gammay ||
      !m.gammaz || (m.nspec_irregular > 0 && !m.irregular_element_number) ||
      (m.nabs > 0 && (!m.abs_ispec || !m.abs_ijk || !m.abs_jacobian2Dw || !m.abs_normal)))
    throw std::invalid_argument("null 3-D database array");
  if (m.nspec_irregular < 0 || m.nspec_irregular > m.nspec || m.nabs < 0)
    throw std::invalid_argument("bad 3-D database counts");
  const int64_t pad64 = (static_cast<int64_t>(m.nglob) + 63) / 64 * 64;
  check_index_range_3d(pad64, m.nspec);
  const int nspec = m.nspec, npe = 125;
  const size_t nloc = static_cast<size_t>(nspec) * npe;
  for (int e = 0; e < nspec; e++)
    if (m.ispec_type[e] != 1) throw std::invalid_argument("3-D database meshes: elastic elements only");
  HostAssembly a;
  a.dim = 3;
  a.ncomp = 3;
  a.npe = npe;
  a.general3d = true;
  a.nspec = nspec;
  a.nspec_el = nspec;
  a.nspec_none = nspec;
  const Gll g = make_gll();
  std::copy(g.xi, g.xi + 5, a.xi);
  std::copy(g.w, g.w + 5, a.w);
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) a.hprime[i][j] = g.hprime[i][j];
  // numbering as the database holds it (generate_databases' ibool)
  a.index_mapping.assign(m.ibool, m.ibool + nloc);
  for (int32_t v : a.index_mapping)
    if (v < 0 || v >= m.nglob) throw std::invalid_argument("ibool out of range");
  a.nglob = a.nglob_el = m.nglob;
  a.nglob_pad = static_cast<int>(pad64);
  a.field_index.resize(m.nglob);
  for (int i = 0; i < m.nglob; i++) a.field_index[i] = i;
  a.gather = a.index_mapping;
  // geometry: per point the nine derivatives and the jacobian, (nspec, 10,
  // 125); regular elements take xix_regular and jacobian_regular
  // (mesh.cpp:150-151)
  std::vector<char> irregular(nspec, 0);
  for (int i = 0; i < m.nspec_irregular; i++) {
    const int e = m.irregular_element_number[i];
    if (e < 0 || e >= nspec) throw std::invalid_argument("irregular element number out of range");
    irregular[e] = 1;
  }
  a.geo3.assign(nloc * 10, 0.0f);
  const float *d9[9] = {m.xix, m.xiy, m.xiz, m.etax, m.etay, m.etaz, m.gammax, m.gammay, m.gammaz};
  for (int e = 0; e < nspec; e++)
    for (int p = 0; p < npe; p++) {
      const size_t q = static_cast<size_t>(e) * npe + p;
      float *gq = &a.geo3[static_cast<size_t>(e) * 10 * npe + p];
      if (irregular[e]) {
        for (int d = 0; d < 9; d++) gq[d * npe] = d9[d][q];
        gq[9 * npe] = m.jacobian[q];
      } else {
        gq[0 * npe] = m.xix_regular;
        gq[4 * npe] = m.xix_regular;
        gq[8 * npe] = m.xix_regular;
        gq[9 * npe] = m.jacobian_regular;
      }
      if (!(gq[9 * npe] > 0.0f))
        throw std::runtime_error("element " + std::to_string(e) + " has a non-positive jacobian");
    }
  // properties per point (mesh.cpp:208-209, 265)
  a.kappa.assign(m.kappa, m.kappa + nloc);
  a.mu.assign(m.mu, m.mu + nloc);
  a.rho.assign(m.rho, m.rho + nloc);
  for (size_t q = 0; q < nloc; q++)
    if (!(a.rho[q] > 0.0f) || !(a.mu[q] > 0.0f) || !(a.kappa[q] + (4.0f / 3.0f) * a.mu[q] > 0.0f))
      throw std::invalid_argument("properties must give rho > 0, mu > 0 and lambda + 2 mu > 0");
  // Stacey: one entry per point of every absorbing face, its unit normal and
  // jacobian2Dw (the traction t = -[rho vp (v.n) n + rho vs (v - (v.n) n)] jacobian2Dw
  // is the stacey.tpp:36-72 form with |n| = 1 and weight jacobian2Dw);
  // entries of one element point are grouped in face order
  for (int f = 0; f < m.nabs; f++) {
    const int e = m.abs_ispec[f];
    if (e < 0 || e >= nspec) throw std::invalid_argument("absorbing face element out of range");
    for (int qf = 0; qf < 25; qf++) {
      const int i = m.abs_ijk[(static_cast<size_t>(f) * 3 + 0) * 25 + qf];
      const int j = m.abs_ijk[(static_cast<size_t>(f) * 3 + 1) * 25 + qf];
      const int k = m.abs_ijk[(static_cast<size_t>(f) * 3 + 2) * 25 + qf];
      if (i < 0 || i > 4 || j < 0 || j > 4 || k < 0 || k > 4) throw std::invalid_argument("face ijk out of range");
      const size_t p = static_cast<size_t>(e) * npe + (k * 5 + j) * 5 + i;
      a.bnd_gather.push_back(a.gather[p]);
      a.bnd_slot.push_back(static_cast<int32_t>(p));
      for (int d = 0; d < 3; d++) a.bnd_normal.push_back(m.abs_normal[(static_cast<size_t>(f) * 3 + d) * 25 + qf]);
      a.bnd_weight.push_back(m.abs_jacobian2Dw[static_cast<size_t>(f) * 25 + qf]);
      a.bnd_rho_vp.push_back(m.rho_vp[p]);
      a.bnd_rho_vs.push_back(m.rho_vs[p]);
    }
  }
  // mass: rho w_i w_j w_k J assembled, plus the Stacey term of a constant
  // velocity -dt/2
  {
    std::vector<float> mass(3 * static_cast<size_t>(a.nglob_pad), 0.0f);
    for (int e = 0; e < nspec; e++)
      for (int k = 0; k < 5; k++)
        for (int j = 0; j < 5; j++)
          for (int i = 0; i < 5; i++) {
            const int p = (k * 5 + j) * 5 + i;
            const size_t q = static_cast<size_t>(e) * npe + p;
            const float f = a.w[i] * a.w[j] * a.w[k] * a.geo3[static_cast<size_t>(e) * 10 * npe + 9 * npe + p];
            const float mm = a.rho[q] * f;
            for (int c = 0; c < 3; c++) mass[static_cast<size_t>(c) * a.nglob_pad + a.gather[q]] += mm;
          }
    const float vfac = -1.0f * static_cast<float>(run.dt) * 0.5f;
    for (size_t q = 0; q < a.bnd_gather.size(); q++) {
      float t[3] = {0, 0, 0};
      stacey_term<3>(&a.bnd_normal[3 * q], a.bnd_weight[q], a.bnd_rho_vp[q], a.bnd_rho_vs[q], vfac, t);
      for (int c = 0; c < 3; c++) mass[static_cast<size_t>(c) * a.nglob_pad + a.bnd_gather[q]] += t[c];
    }
    a.mass_inverse.assign(mass.size(), 0.0f);
    for (size_t i = 0; i < mass.size(); i++)
      if (mass[i] != 0.0f) a.mass_inverse[i] = 1.0f / mass[i];
  }
  // sources and receivers
  Locator3D loc(m, a.xi);
  setup_time_function(a, nsrc, src, run);
  a.nsrc = nsrc;
  a.src_ispec.resize(nsrc);
  a.src_array.assign(static_cast<size_t>(nsrc) * 3 * npe, 0.0f);
  a.src_local.resize(3 * static_cast<size_t>(nsrc));
  for (int s = 0; s < nsrc; s++) {
    int e;
    double ref[3];
    loc.locate(src[s].x, src[s].y, src[s].z, e, ref);
    a.src_ispec[s] = e;
    for (int d = 0; d < 3; d++) a.src_local[3 * s + d] = ref[d];
    float hx[5], hpx[5], hy[5], hpy[5], hz[5], hpz[5];
    lagrange<float>(static_cast<float>(ref[0]), a.xi, hx, hpx);
    lagrange<float>(static_cast<float>(ref[1]), a.xi, hy, hpy);
    lagrange<float>(static_cast<float>(ref[2]), a.xi, hz, hpz);
    // element derivatives averaged with the Lagrange weights of the source
    // point (the 3-D form of moment_tensor_source.cpp:88-130)
    float dav[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++)
        for (int i = 0; i < 5; i++) {
          const int p = (k * 5 + j) * 5 + i;
          const float hl = hx[i] * hy[j] * hz[k];
          for (int d = 0; d < 9; d++) dav[d] += hl * a.geo3[static_cast<size_t>(e) * 10 * npe + d * npe + p];
        }
    float *arr = &a.src_array[static_cast<size_t>(s) * 3 * npe];
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++)
        for (int i = 0; i < 5; i++) {
          const int p = (k * 5 + j) * 5 + i;
          if (src[s].kind == SFX_SOURCE_FORCE) {
            const float h = hx[i] * hy[j] * hz[k];
            arr[p] = static_cast<float>(src[s].fx) * h;
            arr[npe + p] = static_cast<float>(src[s].fy) * h;
            arr[2 * npe + p] = static_cast<float>(src[s].fz) * h;
          } else if (src[s].kind == SFX_SOURCE_MOMENT_TENSOR) {
            const float dref[3] = {hpx[i] * hy[j] * hz[k], hx[i] * hpy[j] * hz[k], hx[i] * hy[j] * hpz[k]};
            float dxr[3];
            for (int r = 0; r < 3; r++) dxr[r] = dref[0] * dav[0 + r] + dref[1] * dav[3 + r] + dref[2] * dav[6 + r];
            const float M[3][3] = {
                {static_cast<float>(src[s].m[0]), static_cast<float>(src[s].m[3]), static_cast<float>(src[s].m[4])},
                {static_cast<float>(src[s].m[3]), static_cast<float>(src[s].m[1]), static_cast<float>(src[s].m[5])},
                {static_cast<float>(src[s].m[4]), static_cast<float>(src[s].m[5]), static_cast<float>(src[s].m[2])}};
            for (int c = 0; c < 3; c++) arr[c * npe + p] = M[c][0] * dxr[0] + M[c][1] * dxr[1] + M[c][2] * dxr[2];
          } else {
            throw std::invalid_argument("unknown source kind");
          }
        }
  }
  for (int s = 0; s < nsrc; s++) a.src_order.push_back(s);
  a.src_order_el = a.src_order;
  a.nrec = nrec;
  a.rec_ispec.resize(nrec);
  a.rec_h.resize(static_cast<size_t>(nrec) * npe);
  a.rec_local.resize(3 * static_cast<size_t>(nrec));
  for (int r = 0; r < nrec; r++) {
    int e;
    double ref[3];
    loc.locate(rec[r].x, rec[r].y, rec[r].z, e, ref);
    a.rec_ispec[r] = e;
    for (int d = 0; d < 3; d++) a.rec_local[3 * r + d] = ref[d];
    float hx[5], hy[5], hz[5];
    lagrange<float>(static_cast<float>(ref[0]), a.xi, hx, nullptr);
    lagrange<float>(static_cast<float>(ref[1]), a.xi, hy, nullptr);
    lagrange<float>(static_cast<float>(ref[2]), a.xi, hz, nullptr);
    for (int k = 0; k < 5; k++)
      for (int j = 0; j < 5; j++)
        for (int i = 0; i < 5; i++) a.rec_h[static_cast<size_t>(r) * npe + (k * 5 + j) * 5 + i] = hx[i] * hy[j] * hz[k];
  }
  return a;
}

void color_elements(HostAssembly &a, int scatter) {
  ElementColoring c = element_coloring(a, scatter);
  a.perm = std::move(c.perm);
  a.inv_perm = std::move(c.inv_perm);
  a.color_offsets = std::move(c.color_offsets);
  a.ncolors_el = c.ncolors_el;
}

ElementColoring element_coloring(const HostAssembly &a, int scatter) {
  ElementColoring out;
  const int nspec = a.nspec, npe = a.npe;
  std::vector<int> color(nspec, 0);
  int ncolors = 1;
  if (scatter == SFX_SCATTER_COLORED) {  // ATOMIC, GATHER and BLOCK keep compute order
    if (a.dim == 3 && !a.general3d) {
      // structured hexahedra: the parity of (ex, ey, ez) is a valid 8-coloring
      // (it is also what first-fit greedy produces in element order)
      for (int e = 0; e < nspec; e++) {
        const int ex = e % a.nx, ey = (e / a.nx) % a.ny, ez = e / (a.nx * a.ny);
        color[e] = (ex & 1) | ((ey & 1) << 1) | ((ez & 1) << 2);
      }
      ncolors = 8;
    } else {
      // first-fit greedy in compute order over the point-sharing graph
      std::vector<int> start(a.nglob + 1, 0);
      for (size_t p = 0; p < a.index_mapping.size(); p++) start[a.index_mapping[p] + 1]++;
      for (int g = 0; g < a.nglob; g++) start[g + 1] += start[g];
      std::vector<int> fill(start.begin(), start.end() - 1), elems(a.index_mapping.size());
      for (int e = 0; e < nspec; e++)
        for (int p = 0; p < npe; p++) elems[fill[a.index_mapping[static_cast<size_t>(e) * npe + p]]++] = e;
      std::fill(color.begin(), color.end(), -1);
      ncolors = 0;
      std::vector<int> stamp(64, -1);
      // elastic and acoustic elements update different fields: each medium is
      // colored on its own, the acoustic colors numbered after the elastic ones
      int base = 0;
      for (int e = 0; e < nspec; e++) {
        if (e == a.nspec_el) base = ncolors;
        const bool ac = e >= a.nspec_el;
        for (int p = 0; p < npe; p++) {
          const int g = a.index_mapping[static_cast<size_t>(e) * npe + p];
          for (int k = start[g]; k < start[g + 1]; k++) {
            const int c = color[elems[k]];
            if (c >= 0 && (elems[k] >= a.nspec_el) == ac) stamp[c - base] = e;
          }
        }
        int c = 0;
        while (c < 64 && stamp[c] == e) c++;
        if (c == 64) throw std::runtime_error("coloring needs more than 64 colors");
        color[e] = base + c;
        ncolors = std::max(ncolors, base + c + 1);
      }
      if (a.coupled() && base == 0 && a.nspec_el > 0) throw std::logic_error("internal: acoustic color base");
    }
  }
  out.ncolors_el = ncolors;
  if (a.coupled()) {
    if (scatter != SFX_SCATTER_COLORED) {  // one atomic launch per medium
      for (int e = 0; e < nspec; e++) color[e] = e >= a.nspec_el ? 1 : 0;
      ncolors = 2;
    }
    out.ncolors_el = 0;
    for (int e = 0; e < a.nspec_el; e++) out.ncolors_el = std::max(out.ncolors_el, color[e] + 1);
    for (int e = a.nspec_el; e < nspec; e++)
      if (color[e] < out.ncolors_el) throw std::logic_error("internal: acoustic element in an elastic color");
  }
  // stable counting sort by color
  out.color_offsets.assign(ncolors + 1, 0);
  for (int e = 0; e < nspec; e++) out.color_offsets[color[e] + 1]++;
  for (int c = 0; c < ncolors; c++) out.color_offsets[c + 1] += out.color_offsets[c];
  std::vector<int> pos(out.color_offsets.begin(), out.color_offsets.end() - 1);
  out.perm.resize(nspec);
  out.inv_perm.resize(nspec);
  for (int e = 0; e < nspec; e++) {
    const int s = pos[color[e]]++;
    out.perm[s] = e;
    out.inv_perm[e] = s;
  }
  return out;
}

}  // namespace sfx
