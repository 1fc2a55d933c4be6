// ORACLE — test infrastructure only.
//
// CPU restatement of the SPECFEM++ 2-D elastic P-SV (isotropic) Newmark path,
// used by tests/ and by bench.py's cpu_baseline leg as the checker.  It is
// never linked into, loaded by or called from the product library.
//
// Every function follows the reference file:line it names.  The arithmetic
// type T is the reference's `type_real` (float for the FP32 oracle, double for
// the FP64 oracle, as with ENABLE_DOUBLE_PRECISION).  With one thread the loop
// order is the reference's Serial-backend order (chunk size 1), including the
// order in which contributions are summed into a global point.
//
// Pinning (checks that hold this restatement to the reference):
//   - GLL nodes/weights/hprime: the known answers of tests/unit-tests/gll,
//     quadrature exactness to degree 5, hprime on constants and linears
//     (tests/test_oracle_fixtures.py::test_gll_known_answers).
//   - geometry: SPECFEM2D xix/xiz/gammax/gammaz/jacobian fixtures of
//     tests/unit-tests/compute/partial_derivatives
//     (tests/test_oracle_fixtures.py::test_partial_derivatives_match_specfem2d).
//   - locate_point: the known answers of tests/unit-tests/algorithms/locate.cpp
//     (tests/test_oracle_fixtures.py::test_locate_point_known_answers).
//   - mass matrix: SPECFEM2D rmass_inverse fixture, rel. L1 <= 1e-5
//     (tests/unit-tests/domain/rmass_inverse_tests.cpp:229-234).
//   - full forward run: SPECFEM2D traces, rel. L1 <= 1e-3
//     (tests/unit-tests/displacement_tests/Newmark/newmark_tests.cpp:325-346),
//     free-surface and Stacey fixtures.
//   - numbering order: hand-derived known answers on 2x2-element meshes at
//     storage chunk 1 and 2 (tests/test_numbering_known_answers.py).
//   - stiffness: dense brute-force weak-form assembly on small meshes
//     (symmetry, rigid-body null space, match of K u;
//     tests/test_oracle_bruteforce.py).
//   - pressure seismogram: closed form on a linear displacement field
//     (tests/test_oracle_seismograms.py).
//   - adjoint loop and Frechet kernels: time reversibility and central finite
//     differences of the FP64 misfit (tests/test_oracle_adjoint.py).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

constexpr int NGLL = 5;

// ---------------------------------------------------------------------------
// Quadrature: src/quadrature/gll/gll_library.cpp:11-33 (pnleg),
// src/quadrature/gll/gll.cpp:56-62, lagrange_poly.cpp:12-43 (interpolants),
// lagrange_poly.cpp:161-191 (hprime(i,j) = l_j'(xi_i)).
// ---------------------------------------------------------------------------
template <class T> T pnleg(T z, int n) {
  T p1 = 1.0, p2 = z, p3 = p2;
  for (int k = 1; k < n; k++) {
    double dk = static_cast<double>(k);
    p3 = ((2.0 * dk + 1.0) * z * p2 - dk * p1) / (dk + 1.0);
    p1 = p2;
    p2 = p3;
  }
  return p3;
}

template <class T> T pndleg(T z, int n) {
  // derivative of the Legendre polynomial (gll_library.cpp:53-78 restated by
  // the recurrence P'_{k+1} = (2k+1) P_k + P'_{k-1}).
  T p1 = 1.0, p2 = z, p1d = 0.0, p2d = 1.0, p3d = 1.0;
  for (int k = 1; k < n; k++) {
    double dk = static_cast<double>(k);
    T p3 = ((2.0 * dk + 1.0) * z * p2 - dk * p1) / (dk + 1.0);
    p3d = ((2.0 * dk + 1.0) * p2 + (2.0 * dk + 1.0) * z * p2d - dk * p1d) / (dk + 1.0);
    p1 = p2;
    p2 = p3;
    p1d = p2d;
    p2d = p3d;
  }
  return p3d;
}

template <class T> struct Quadrature {
  T xi[NGLL], w[NGLL], hprime[NGLL][NGLL];  // hprime[i][j] = l_j'(xi_i)
  T hprime_wgll[NGLL][NGLL];                // [a][b] = hprime[b][a] * w[b]
};

// GLL nodes and weights for alpha = beta = 0: interior nodes are the roots of
// P'_{N-1}; weights 2 / (N (N-1) P_{N-1}(x)^2).  Computed in double by Newton
// iteration and rounded to T (the reference's zwgljd, gll_library.cpp:102-148,
// converges to the same roots).
template <class T> Quadrature<T> make_quadrature() {
  Quadrature<T> q;
  const int n = NGLL - 1;
  double x[NGLL];
  x[0] = -1.0;
  x[n] = 1.0;
  for (int i = 1; i < n; i++) {
    double z = -std::cos(M_PI * i / n);  // Chebyshev-Gauss-Lobatto guess
    for (int it = 0; it < 100; it++) {
      // Newton on P'_n using P''_n from the Legendre ODE:
      // (1 - z^2) P'' = 2 z P' - n (n+1) P
      double p = pnleg<double>(z, n), dp = pndleg<double>(z, n);
      double d2p = (2.0 * z * dp - n * (n + 1.0) * p) / (1.0 - z * z);
      double dz = dp / d2p;
      z -= dz;
      if (std::abs(dz) < 1e-17) break;
    }
    x[i] = z;
  }
  if (NGLL % 2) x[n / 2] = 0.0;
  for (int i = 0; i < NGLL; i++) {
    double p = pnleg<double>(x[i], n);
    q.xi[i] = static_cast<T>(x[i]);
    q.w[i] = static_cast<T>(2.0 / (n * (n + 1.0) * p * p));
  }
  const int degpoly = n;
  for (int i = 0; i < NGLL; i++)
    for (int j = 0; j < NGLL; j++) {
      T v;
      if (i == 0 && j == 0)
        v = -1.0 * static_cast<T>(degpoly) * (static_cast<T>(degpoly) + 1.0) * 0.25;
      else if (i == degpoly && j == degpoly)
        v = 1.0 * static_cast<T>(degpoly) * (static_cast<T>(degpoly) + 1.0) * 0.25;
      else if (i == j)
        v = 0.0;
      else
        v = pnleg<T>(q.xi[i], degpoly) /
                (pnleg<T>(q.xi[j], degpoly) * (q.xi[i] - q.xi[j])) +
            (1.0 - q.xi[i] * q.xi[i]) * pndleg<T>(q.xi[i], degpoly) /
                (static_cast<T>(degpoly) * (static_cast<T>(degpoly) + 1.0) *
                 pnleg<T>(q.xi[j], degpoly) * (q.xi[i] - q.xi[j]) * (q.xi[i] - q.xi[j]));
      q.hprime[i][j] = v;
    }
  // include/compute/compute_mesh.hpp:266-272
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) q.hprime_wgll[ix][iz] = q.hprime[iz][ix] * q.w[iz];
  return q;
}

// lagrange_poly.cpp:12-43
template <class T> void lagrange_interpolants(T xi, const T *xigll, T *h, T *hprime) {
  for (int dgr = 0; dgr < NGLL; dgr++) {
    T prod1 = 1.0, prod2 = 1.0, prod3, sum = 0.0;
    T x0 = xigll[dgr];
    for (int i = 0; i < NGLL; i++) {
      if (i != dgr) {
        T x = xigll[i];
        prod1 = prod1 * (xi - x);
        prod2 = prod2 * (x0 - x);
        prod3 = 1.0;
        for (int j = 0; j < NGLL; j++)
          if (j != dgr && j != i) prod3 = prod3 * (xi - xigll[j]);
        sum = sum + prod3;
      }
    }
    T prod2_inv = 1.0 / prod2;
    h[dgr] = prod1 * prod2_inv;
    if (hprime) hprime[dgr] = sum * prod2_inv;
  }
}

// ---------------------------------------------------------------------------
// Shape functions: src/jacobian/shape_functions.cpp:17-152
// ---------------------------------------------------------------------------
template <class T> void shape2d(T xi, T gamma, int ngnod, T *s) {
  const T one = 1.0, two = 2.0, quarter = 0.25, half = 0.5;
  const T sp = xi + one, sm = xi - one, tp = gamma + one, tm = gamma - one;
  if (ngnod == 4) {
    s[0] = quarter * sm * tm;
    s[1] = -quarter * sp * tm;
    s[2] = quarter * sp * tp;
    s[3] = -quarter * sm * tp;
  } else if (ngnod == 9) {
    const T ss = xi * xi, tt = gamma * gamma, st = xi * gamma;
    (void)two;
    s[0] = quarter * sm * st * tm;
    s[1] = quarter * sp * st * tm;
    s[2] = quarter * sp * st * tp;
    s[3] = quarter * sm * st * tp;
    s[4] = half * tm * gamma * (one - ss);
    s[5] = half * sp * xi * (one - tt);
    s[6] = half * tp * gamma * (one - ss);
    s[7] = half * sm * xi * (one - tt);
    s[8] = (one - ss) * (one - tt);
  } else {
    throw std::invalid_argument("wrong number of control nodes");
  }
}

template <class T> void dshape2d(T xi, T gamma, int ngnod, T *d0, T *d1) {
  const T one = 1.0, two = 2.0, quarter = 0.25, half = 0.5;
  const T sp = xi + one, sm = xi - one, tp = gamma + one, tm = gamma - one;
  if (ngnod == 4) {
    d0[0] = quarter * tm;
    d0[1] = -quarter * tm;
    d0[2] = quarter * tp;
    d0[3] = -quarter * tp;
    d1[0] = quarter * sm;
    d1[1] = -quarter * sp;
    d1[2] = quarter * sp;
    d1[3] = -quarter * sm;
  } else if (ngnod == 9) {
    const T s2 = xi * two, t2 = gamma * two, ss = xi * xi, tt = gamma * gamma, st = xi * gamma;
    d0[0] = quarter * tm * gamma * (s2 - one);
    d0[1] = quarter * tm * gamma * (s2 + one);
    d0[2] = quarter * tp * gamma * (s2 + one);
    d0[3] = quarter * tp * gamma * (s2 - one);
    d1[0] = quarter * sm * xi * (t2 - one);
    d1[1] = quarter * sp * xi * (t2 - one);
    d1[2] = quarter * sp * xi * (t2 + one);
    d1[3] = quarter * sm * xi * (t2 + one);
    d0[4] = -one * st * tm;
    d0[5] = half * (one - tt) * (s2 + one);
    d0[6] = -one * st * tp;
    d0[7] = half * (one - tt) * (s2 - one);
    d1[4] = half * (one - ss) * (t2 - one);
    d1[5] = -one * st * sp;
    d1[6] = half * (one - ss) * (t2 + one);
    d1[7] = -one * st * sm;
    d0[8] = -one * s2 * (one - tt);
    d1[8] = -one * t2 * (one - ss);
  } else {
    throw std::invalid_argument("wrong number of control nodes");
  }
}

// ---------------------------------------------------------------------------
// Mesh input (mesh order) and assembled state (compute order)
// ---------------------------------------------------------------------------
enum Edge { BOTTOM = 0, RIGHT = 1, TOP = 2, LEFT = 3, BOTTOM_LEFT = 4, BOTTOM_RIGHT = 5, TOP_RIGHT = 6, TOP_LEFT = 7 };

// Element boundary tags (include/enumerations/boundary.hpp: the container's
// += combines a Stacey and an acoustic free-surface tag into the composite).
enum BTag { BT_NONE = 0, BT_AFS = 1, BT_STACEY = 2, BT_COMPOSITE = 3 };
inline int add_tag(int cur, int t) {
  if (cur == BT_NONE) return t;
  if (cur == t || cur == BT_COMPOSITE) return cur;
  return BT_COMPOSITE;  // Stacey + free surface in either order
}
// Point on an edge or corner of an element
// (src/compute/boundaries/impl/acoustic_free_surface.cpp:10-26).
inline bool on_boundary(int type, int iz, int ix) {
  const int n = NGLL - 1;
  return (type == TOP && iz == n) || (type == BOTTOM && iz == 0) || (type == LEFT && ix == 0) ||
         (type == RIGHT && ix == n) || (type == BOTTOM_RIGHT && iz == 0 && ix == n) ||
         (type == BOTTOM_LEFT && iz == 0 && ix == 0) || (type == TOP_RIGHT && iz == n && ix == n) ||
         (type == TOP_LEFT && iz == n && ix == 0);
}

struct MeshIn {
  int nspec, ngnod, npgeo;
  const double *coorg;  // (npgeo, 2)
  const int *knods;     // (nspec, ngnod)
  const int *kmato;     // (nspec)
  int nmat;
  const double *rho, *cp, *cs;  // per material
  int nabs;
  const int *abs_ispec, *abs_type;  // mesh element and Edge of each absorbing edge
  int nafs = 0;                     // acoustic free-surface entries: element, Edge (edge or corner)
  const int *afs_ispec = nullptr, *afs_type = nullptr;
  int nfs = 0;                      // fluid-solid edges: acoustic and elastic mesh element
  const int *fs_acoustic = nullptr, *fs_elastic = nullptr;
};

// One fluid-solid edge of the elastic-acoustic interface container
// (include/compute/coupled_interfaces/interface_container.tpp): elements in
// compute order, the connected edge of each, and per edge point the
// integration factor and the elastic element's edge normal.
template <class T> struct CoupledEdge {
  int e_el, e_ac, edge_el, edge_ac;
  T factor[NGLL], nx[NGLL], nz[NGLL];
};

template <class T> struct Assembly {
  int nspec = 0, ngnod = 0, nglob = 0, chunk = 1;
  Quadrature<T> q;
  std::vector<int> compute_to_mesh, mesh_to_compute;
  std::vector<int> is_stacey;        // per compute element
  int nspec_none = 0;                // elements [0, nspec_none) have no BC
  std::vector<T> ctrl;               // (ispec, in, 2) control node coordinates
  std::vector<int> ibool;            // (ispec, iz, ix) global index
  std::vector<int> medium_index;     // iglob -> per-medium index
  std::vector<T> coord;              // (ispec, iz, ix, 2) point coordinates
  std::vector<T> xix, xiz, gammax, gammaz, jac;  // (ispec, iz, ix)
  std::vector<T> kappa, mu, rho;     // (ispec, iz, ix)
  // Stacey: per point of a Stacey element; tag 1 if the point is on an
  // absorbing edge (compute/boundaries/impl/stacey.cpp:140-180)
  std::vector<int> bnd_tag;          // (ispec, iz, ix), 0 for non-Stacey elements
  std::vector<T> bnd_nx, bnd_nz, bnd_w;
  std::vector<T> mass_inverse;       // (comp, nglob) LayoutLeft on medium index
  // acoustic medium (the elastic arrays above then cover the elastic elements;
  // medium_index is -1 at points no elastic element touches)
  std::vector<int> medium;           // per compute element: 0 elastic P-SV, 1 acoustic
  std::vector<int> btag;             // per compute element: BTag
  std::vector<int> dir_tag;          // (ispec, iz, ix): 1 on an acoustic free surface
  std::vector<int> index_ac;         // iglob -> acoustic field index, or -1
  int nglob_el = 0, nglob_ac = 0;
  std::vector<T> rho_inv;            // acoustic: 1/rho per point
  std::vector<T> mass_inverse_ac;    // (nglob_ac)
  std::vector<CoupledEdge<T>> coupling;
  bool coupled() const { return nglob_ac > 0; }
  double dt = 0;                     // time step of the Stacey mass term
  double start_time = 0;             // user start time (0: automatic), see make_sources
};

template <class T> void compute_mass(Assembly<T> &a);

// src/compute/compute_mesh.cpp:19-43 (note: evaluated on the *sorted* points,
// and numeric_limits::min() is the smallest positive value, both as in the
// reference).
template <class T> struct QP {
  T x, z;
  int iloc, iglob;
};

template <class T>
T get_tolerance(const std::vector<QP<T>> &c, int nspec, int ngllxz) {
  T xtypdist = std::numeric_limits<T>::max();
  for (int ispec = 0; ispec < nspec; ispec++) {
    T xmax = std::numeric_limits<T>::min(), xmin = std::numeric_limits<T>::max();
    T ymax = std::numeric_limits<T>::min(), ymin = std::numeric_limits<T>::max();
    for (int xz = 0; xz < ngllxz; xz++) {
      int iloc = ispec * ngllxz + xz;
      xmax = std::max(xmax, c[iloc].x);
      xmin = std::min(xmin, c[iloc].x);
      ymax = std::max(ymax, c[iloc].z);
      ymin = std::min(ymin, c[iloc].z);
    }
    xtypdist = std::min(xtypdist, xmax - xmin);
    xtypdist = std::min(xtypdist, ymax - ymin);
  }
  return 1e-6 * xtypdist;
}

// src/compute/compute_mesh.cpp:45-170 (assign_numbering) with a selectable
// storage chunk size (include/parallel_configuration/chunk_config.hpp:12-28).
template <class T>
void assign_numbering(Assembly<T> &a, const std::vector<double> &gcoord /*(ispec,iz,ix,2)*/) {
  const int nspec = a.nspec, ngll = NGLL, ngllxz = NGLL * NGLL, chunk = a.chunk;
  std::vector<QP<T>> cc(static_cast<size_t>(nspec) * ngllxz);
  int iloc = 0;
  for (int ichunk = 0; ichunk < nspec; ichunk += chunk)
    for (int iz = 0; iz < ngll; iz++)
      for (int ix = 0; ix < ngll; ix++)
        for (int ie = 0; ie < chunk; ie++) {
          int ispec = ichunk + ie;
          if (ispec >= nspec) break;
          size_t g = ((static_cast<size_t>(ispec) * ngll + iz) * ngll + ix) * 2;
          cc[iloc].x = static_cast<T>(gcoord[g]);
          cc[iloc].z = static_cast<T>(gcoord[g + 1]);
          cc[iloc].iloc = iloc;
          cc[iloc].iglob = 0;
          iloc++;
        }
  std::sort(cc.begin(), cc.end(), [](const QP<T> &p, const QP<T> &q) {
    if (p.x != q.x) return p.x < q.x;
    return p.z < q.z;
  });
  int ig = 0;
  cc[0].iglob = ig;
  T xtol = get_tolerance(cc, nspec, ngllxz);
  for (size_t i = 1; i < cc.size(); i++) {
    if ((std::abs(cc[i].x - cc[i - 1].x) > xtol) || (std::abs(cc[i].z - cc[i - 1].z) > xtol)) ig++;
    cc[i].iglob = ig;
  }
  std::vector<QP<T>> orig(cc.size());
  for (size_t i = 0; i < cc.size(); i++) orig[cc[i].iloc] = cc[i];
  const int nglob = ig + 1;
  std::vector<int> counted(nglob, -1);
  a.ibool.assign(static_cast<size_t>(nspec) * ngllxz, -1);
  a.coord.assign(static_cast<size_t>(nspec) * ngllxz * 2, 0);
  iloc = 0;
  int inum = 0;
  for (int ichunk = 0; ichunk < nspec; ichunk += chunk)
    for (int iz = 0; iz < ngll; iz++)
      for (int ix = 0; ix < ngll; ix++)
        for (int ie = 0; ie < chunk; ie++) {
          int ispec = ichunk + ie;
          if (ispec >= nspec) break;
          size_t p = (static_cast<size_t>(ispec) * ngll + iz) * ngll + ix;
          if (counted[orig[iloc].iglob] == -1) {
            counted[orig[iloc].iglob] = inum;
            a.ibool[p] = inum++;
          } else {
            a.ibool[p] = counted[orig[iloc].iglob];
          }
          a.coord[2 * p] = orig[iloc].x;
          a.coord[2 * p + 1] = orig[iloc].z;
          iloc++;
        }
  if (inum != nglob) throw std::runtime_error("numbering: count mismatch");
  a.nglob = nglob;
}

// Per-medium numbering, include/compute/fields/impl/field_impl.tpp:39-62
// (loop order chunk, ix, iz, element; elements of other media skipped).
template <class T> int number_medium(const Assembly<T> &a, int med, std::vector<int> &index) {
  index.assign(a.nglob, -1);
  int count = 0;
  for (int ichunk = 0; ichunk < a.nspec; ichunk += a.chunk)
    for (int ix = 0; ix < NGLL; ix++)
      for (int iz = 0; iz < NGLL; iz++)
        for (int ie = 0; ie < a.chunk; ie++) {
          int ispec = ichunk + ie;
          if (ispec >= a.nspec) break;
          if (a.medium[ispec] != med) continue;
          int g = a.ibool[(static_cast<size_t>(ispec) * NGLL + iz) * NGLL + ix];
          if (index[g] == -1) index[g] = count++;
        }
  return count;
}
template <class T> void medium_numbering(Assembly<T> &a) {
  a.nglob_el = number_medium(a, 0, a.medium_index);
  a.nglob_ac = number_medium(a, 1, a.index_ac);
  if (a.nglob_ac == 0 && a.nglob_el != a.nglob) throw std::runtime_error("medium numbering: count mismatch");
}

template <class T> inline size_t pidx(int ispec, int iz, int ix) {
  return (static_cast<size_t>(ispec) * NGLL + iz) * NGLL + ix;
}

// Stacey point term, include/boundary_conditions/stacey/stacey.tpp:36-72.
template <class T>
inline void stacey_traction(T nx, T nz, T w, T kappa, T mu, T rho, T vx, T vz, T &tx, T &tz) {
  const T vn = vx * nx + vz * nz;
  const T jacobian1d = std::sqrt(nx * nx + nz * nz);
  const T lp2m = kappa + static_cast<T>(4.0 / 3.0) * mu;
  const T rho_vp = std::sqrt(lp2m / rho) * rho;
  const T rho_vs = std::sqrt(rho * mu);
  const T f0 = ((vn * nx / (jacobian1d * jacobian1d)) * (rho_vp - rho_vs)) + vx * rho_vs;
  const T f1 = ((vn * nz / (jacobian1d * jacobian1d)) * (rho_vp - rho_vs)) + vz * rho_vs;
  tx += static_cast<T>(-1.0) * f0 * jacobian1d * w;
  tz += static_cast<T>(-1.0) * f1 * jacobian1d * w;
}

// Points of an edge in increasing order, (ix, iz) (interface_container.tpp:15-48).
inline void edge_point(int edge, int i, int &ix, int &iz) {
  switch (edge) {
    case BOTTOM: ix = i; iz = 0; break;
    case TOP: ix = i; iz = NGLL - 1; break;
    case LEFT: ix = 0; iz = i; break;
    default: ix = NGLL - 1; iz = i; break;  // RIGHT
  }
}

// Edge normal scaled by the Jacobian (core/specfem/point/partial_derivatives.hpp:371-396).
template <class T> void edge_normal(const Assembly<T> &a, size_t p, int edge, T &nx, T &nz) {
  const T J = a.jac[p];
  switch (edge) {
    case BOTTOM: nx = static_cast<T>(-1.0) * a.gammax[p] * J; nz = static_cast<T>(-1.0) * a.gammaz[p] * J; break;
    case TOP: nx = a.gammax[p] * J; nz = a.gammaz[p] * J; break;
    case LEFT: nx = static_cast<T>(-1.0) * a.xix[p] * J; nz = static_cast<T>(-1.0) * a.xiz[p] * J; break;
    default: nx = a.xix[p] * J; nz = a.xiz[p] * J; break;  // RIGHT
  }
}

// Elastic-acoustic interface container
// (include/compute/coupled_interfaces/interface_container.tpp:50-365):
// the one pair of edges of the two elements whose points coincide (relative
// distance <= 1e-10 of the first element's edge spacing), the GLL weight
// along the edge as the factor and the elastic element's edge normals.
template <class T> void build_coupling(Assembly<T> &a, const MeshIn &m) {
  const int edges[4] = {BOTTOM, TOP, LEFT, RIGHT};
  a.coupling.clear();
  for (int i = 0; i < m.nfs; i++) {
    CoupledEdge<T> ce{};
    ce.e_el = a.mesh_to_compute[m.fs_elastic[i]];
    ce.e_ac = a.mesh_to_compute[m.fs_acoustic[i]];
    if (!((a.medium[ce.e_el] == 0 && a.medium[ce.e_ac] == 1) || (a.medium[ce.e_el] == 1 && a.medium[ce.e_ac] == 0)))
      throw std::runtime_error("Coupled Interfaces: Interface is not between the correct mediums");
    const int e1 = ce.e_el, e2 = ce.e_ac;
    int found = 0;
    for (int ed1 : edges)
      for (int ed2 : edges) {
        if (ed1 == ed2) continue;
        const T edge_size = (ed1 == BOTTOM || ed1 == TOP)
                                ? a.coord[2 * pidx<T>(e1, 0, 1)] - a.coord[2 * pidx<T>(e1, 0, 0)]
                                : a.coord[2 * pidx<T>(e1, 1, 0) + 1] - a.coord[2 * pidx<T>(e1, 0, 0) + 1];
        bool connected = true;
        for (int ip = 0; ip < NGLL && connected; ip++) {
          int ix1, iz1, ix2, iz2;
          edge_point(ed1, ip, ix1, iz1);
          edge_point(ed2, ip, ix2, iz2);
          const size_t p1 = pidx<T>(e1, iz1, ix1), p2 = pidx<T>(e2, iz2, ix2);
          const T dx = a.coord[2 * p1] - a.coord[2 * p2], dz = a.coord[2 * p1 + 1] - a.coord[2 * p2 + 1];
          const T distance = std::sqrt(dx * dx + dz * dz);
          if (distance / edge_size > 1.e-10) connected = false;
        }
        if (connected) {
          ce.edge_el = ed1;
          ce.edge_ac = ed2;
          found++;
        }
      }
    if (found != 1) throw std::runtime_error("Number of connected edges is not 1");
    for (int ip = 0; ip < NGLL; ip++) {
      int ix1, iz1, ix2, iz2;
      edge_point(ce.edge_el, ip, ix1, iz1);
      edge_point(ce.edge_ac, ip, ix2, iz2);
      T n1x, n1z, n2x, n2z;
      edge_normal(a, pidx<T>(e1, iz1, ix1), ce.edge_el, n1x, n1z);
      edge_normal(a, pidx<T>(e2, iz2, ix2), ce.edge_ac, n2x, n2z);
      if ((std::abs(n1x + n2x) > 1.e-4 * (std::abs(n1x - n2x) / 2)) &&
          (std::abs(n1z + n2z) > 1.e-4 * (std::abs(n1z - n2z) / 2)))
        throw std::runtime_error("Edge normals need to be opposite in direction");
      ce.factor[ip] = (ce.edge_el == BOTTOM || ce.edge_el == TOP) ? a.q.w[ix1] : a.q.w[iz1];
      ce.nx[ip] = n1x;
      ce.nz[ip] = n1z;
    }
    a.coupling.push_back(ce);
  }
}

template <class T> Assembly<T> *build_assembly(const MeshIn &m, int chunk, double dt) {
  auto *a = new Assembly<T>();
  a->nspec = m.nspec;
  a->ngnod = m.ngnod;
  a->chunk = chunk;
  a->q = make_quadrature<T>();
  const int nspec = m.nspec;
  // Materials with vs = 0 are acoustic
  // (src/io/mesh/impl/fortran/dim2/read_material_properties.cpp:172-238).
  std::vector<int> mat_acoustic(m.nmat);
  for (int i = 0; i < m.nmat; i++) mat_acoustic[i] = m.cs[i] == 0.0 ? 1 : 0;
  std::vector<int> medium_mesh(nspec);
  for (int e = 0; e < nspec; e++) medium_mesh[e] = mat_acoustic[m.kmato[e]];

  // Element boundary tags (src/mesh/dim2/tags/tags.cpp:5-47) and compute
  // ordering (src/compute/compute_mesh.cpp:240-279) by the element types of
  // include/enumerations/material_definitions.hpp:76-103: elastic none,
  // elastic Stacey, acoustic none, free surface, Stacey, composite, each in
  // mesh order.
  std::vector<int> tag_mesh(nspec, BT_NONE);
  for (int i = 0; i < m.nabs; i++) tag_mesh[m.abs_ispec[i]] = add_tag(tag_mesh[m.abs_ispec[i]], BT_STACEY);
  for (int i = 0; i < m.nafs; i++) {
    const int e = m.afs_ispec[i];
    if (!medium_mesh[e]) throw std::runtime_error("Acoustic free surface boundary is not an acoustic element");
    tag_mesh[e] = add_tag(tag_mesh[e], BT_AFS);
  }
  for (int e = 0; e < nspec; e++)
    if (!medium_mesh[e] && tag_mesh[e] != BT_NONE && tag_mesh[e] != BT_STACEY)
      throw std::runtime_error("elastic element with a free-surface tag");
  const int group_order[6][2] = {{0, BT_NONE}, {0, BT_STACEY}, {1, BT_NONE}, {1, BT_AFS}, {1, BT_STACEY},
                                 {1, BT_COMPOSITE}};
  a->compute_to_mesh.clear();
  for (auto &grp : group_order)
    for (int e = 0; e < nspec; e++)
      if (medium_mesh[e] == grp[0] && tag_mesh[e] == grp[1]) a->compute_to_mesh.push_back(e);
  a->mesh_to_compute.assign(nspec, 0);
  a->is_stacey.assign(nspec, 0);
  a->medium.assign(nspec, 0);
  a->btag.assign(nspec, 0);
  for (int c = 0; c < nspec; c++) {
    const int e = a->compute_to_mesh[c];
    a->mesh_to_compute[e] = c;
    a->medium[c] = medium_mesh[e];
    a->btag[c] = tag_mesh[e];
    a->is_stacey[c] = (tag_mesh[e] == BT_STACEY || tag_mesh[e] == BT_COMPOSITE) ? 1 : 0;
  }
  a->nspec_none = 0;
  for (int c = 0; c < nspec; c++) a->nspec_none += (a->medium[c] == 0 && !a->is_stacey[c]) ? 1 : 0;

  // Control nodes in compute order (compute_mesh.cpp:172-197), as type_real.
  a->ctrl.resize(static_cast<size_t>(nspec) * m.ngnod * 2);
  for (int c = 0; c < nspec; c++)
    for (int in = 0; in < m.ngnod; in++) {
      int node = m.knods[static_cast<size_t>(a->compute_to_mesh[c]) * m.ngnod + in];
      a->ctrl[(static_cast<size_t>(c) * m.ngnod + in) * 2] = static_cast<T>(m.coorg[2 * node]);
      a->ctrl[(static_cast<size_t>(c) * m.ngnod + in) * 2 + 1] = static_cast<T>(m.coorg[2 * node + 1]);
    }

  // GLL point coordinates (compute_mesh.cpp:299-333): float products summed in
  // double.
  std::vector<double> gcoord(static_cast<size_t>(nspec) * NGLL * NGLL * 2);
  std::vector<T> sh(m.ngnod);
  for (int c = 0; c < nspec; c++)
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        shape2d<T>(a->q.xi[ix], a->q.xi[iz], m.ngnod, sh.data());
        double xc = 0.0, zc = 0.0;
        for (int in = 0; in < m.ngnod; in++) {
          xc += a->ctrl[(static_cast<size_t>(c) * m.ngnod + in) * 2] * sh[in];
          zc += a->ctrl[(static_cast<size_t>(c) * m.ngnod + in) * 2 + 1] * sh[in];
        }
        gcoord[pidx<T>(c, iz, ix) * 2] = xc;
        gcoord[pidx<T>(c, iz, ix) * 2 + 1] = zc;
      }
  assign_numbering(*a, gcoord);
  medium_numbering(*a);

  // Partial derivatives (src/compute/compute_partial_derivatives.cpp:46-91,
  // src/jacobian/jacobian.cpp:232-257).
  const size_t np = static_cast<size_t>(nspec) * NGLL * NGLL;
  a->xix.resize(np); a->xiz.resize(np); a->gammax.resize(np); a->gammaz.resize(np); a->jac.resize(np);
  std::vector<T> d0(m.ngnod), d1(m.ngnod);
  for (int c = 0; c < nspec; c++)
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        dshape2d<T>(a->q.xi[ix], a->q.xi[iz], m.ngnod, d0.data(), d1.data());
        T xxi = 0.0, zxi = 0.0, xgamma = 0.0, zgamma = 0.0;
        for (int in = 0; in < m.ngnod; in++) {
          const T cx = a->ctrl[(static_cast<size_t>(c) * m.ngnod + in) * 2];
          const T cz = a->ctrl[(static_cast<size_t>(c) * m.ngnod + in) * 2 + 1];
          xxi += d0[in] * cx;
          zxi += d0[in] * cz;
          xgamma += d1[in] * cx;
          zgamma += d1[in] * cz;
        }
        const T jacobian = xxi * zgamma - xgamma * zxi;
        if (jacobian < static_cast<T>(1e-10))
          throw std::runtime_error("small jacobian");  // compute_partial_derivatives.cpp:114-150
        size_t p = pidx<T>(c, iz, ix);
        a->xix[p] = zgamma / jacobian;
        a->gammax[p] = -zxi / jacobian;
        a->xiz[p] = -xgamma / jacobian;
        a->gammaz[p] = xxi / jacobian;
        a->jac[p] = jacobian;
      }

  // Properties per point (include/medium/dim2/elastic/isotropic/material.hpp:62-77,126).
  // Acoustic (include/medium/dim2/acoustic/isotropic/material.hpp:44-101):
  // kappa = rho cp^2, rho_inverse = 1 / rho.
  a->kappa.resize(np); a->mu.resize(np); a->rho.resize(np); a->rho_inv.assign(np, 0);
  for (int c = 0; c < nspec; c++) {
    const int mat = m.kmato[a->compute_to_mesh[c]];
    const T density = static_cast<T>(m.rho[mat]);
    const T cs = static_cast<T>(m.cs[mat]);
    const T cp = static_cast<T>(m.cp[mat]);
    const bool ac = a->medium[c] == 1;
    const T mu = ac ? static_cast<T>(0) : density * cs * cs;
    const T kappa = ac ? density * cp * cp : density * (cp * cp - (4.0 / 3.0) * cs * cs);
    const T rinv = ac ? static_cast<T>(1.0) / density : static_cast<T>(0);
    for (int k = 0; k < NGLL * NGLL; k++) {
      a->kappa[static_cast<size_t>(c) * NGLL * NGLL + k] = kappa;
      a->mu[static_cast<size_t>(c) * NGLL * NGLL + k] = mu;
      a->rho[static_cast<size_t>(c) * NGLL * NGLL + k] = density;
      a->rho_inv[static_cast<size_t>(c) * NGLL * NGLL + k] = rinv;
    }
  }

  // Stacey edge data (src/compute/boundaries/impl/stacey.cpp:64-181).  The
  // absorbing-edge list of the mesh never contains corners
  // (src/io/mesh/impl/fortran/dim2/read_boundaries.cpp:10-96 finds none), and
  // at a point shared by two listed edges of one element the last edge wins.
  a->bnd_tag.assign(np, 0);
  a->bnd_nx.assign(np, 0);
  a->bnd_nz.assign(np, 0);
  a->bnd_w.assign(np, 0);
  std::map<int, std::vector<int>> by_elem;
  for (int i = 0; i < m.nabs; i++) by_elem[a->mesh_to_compute[m.abs_ispec[i]]].push_back(i);
  for (auto &kv : by_elem) {
    const int c = kv.first;
    for (int i : kv.second) {
      const int type = m.abs_type[i];
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          bool on = (type == TOP && iz == NGLL - 1) || (type == BOTTOM && iz == 0) ||
                    (type == LEFT && ix == 0) || (type == RIGHT && ix == NGLL - 1);
          if (!on) continue;
          size_t p = pidx<T>(c, iz, ix);
          a->bnd_tag[p] = 1;
          const T J = a->jac[p];
          T nx, nz, w;
          if (type == LEFT) {
            nx = static_cast<T>(-1.0) * a->xix[p] * J;
            nz = static_cast<T>(-1.0) * a->xiz[p] * J;
            w = a->q.w[iz];
          } else if (type == RIGHT) {
            nx = a->xix[p] * J;
            nz = a->xiz[p] * J;
            w = a->q.w[iz];
          } else if (type == TOP) {
            nx = a->gammax[p] * J;
            nz = a->gammaz[p] * J;
            w = a->q.w[ix];
          } else {
            nx = static_cast<T>(-1.0) * a->gammax[p] * J;
            nz = static_cast<T>(-1.0) * a->gammaz[p] * J;
            w = a->q.w[ix];
          }
          a->bnd_nx[p] = nx;
          a->bnd_nz[p] = nz;
          a->bnd_w[p] = w;
        }
    }
  }

  // Acoustic free-surface points (acoustic_free_surface.cpp:28-103).
  a->dir_tag.assign(np, 0);
  for (int i = 0; i < m.nafs; i++) {
    const int c = a->mesh_to_compute[m.afs_ispec[i]];
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++)
        if (on_boundary(m.afs_type[i], iz, ix)) a->dir_tag[pidx<T>(c, iz, ix)] = 1;
  }

  build_coupling(*a, m);
  a->dt = dt;
  compute_mass(*a);
  return a;
}

// Mass matrix (include/kokkos_kernels/impl/compute_mass_matrix.tpp:80-114,
// include/medium/dim2/elastic/isotropic/mass_matrix.tpp:5-16, Stacey term
// stacey.tpp:641-677), then inversion (invert_mass_matrix.tpp:40-47).
// Acoustic Stacey point term (stacey.tpp:172-213): -w |n| rho_vpinverse
// times the velocity, rho_vpinverse = sqrt(rho_inverse / kappa)
// (core/specfem/point/properties.hpp:68-76).
template <class T> inline void stacey_traction_ac(T nx, T nz, T w, T rho_inv, T kappa, T v, T &t) {
  const T factor = w * std::sqrt(nx * nx + nz * nz);
  const T kappa_inverse = static_cast<T>(1.0) / kappa;
  const T rho_vpinverse = std::sqrt(rho_inv * kappa_inverse);
  t += static_cast<T>(-1.0) * factor * rho_vpinverse * v;
}

template <class T> void compute_mass(Assembly<T> &A) {
  Assembly<T> *a = &A;
  const int nspec = a->nspec;
  const double dt = a->dt;
  const int nglob = a->nglob_el;
  std::vector<T> mass(2 * static_cast<size_t>(nglob), 0);
  std::vector<T> mass_ac(static_cast<size_t>(a->nglob_ac), 0);
  const T velocity_factor = static_cast<T>(-1.0) * static_cast<T>(dt) * static_cast<T>(0.5);
  for (int c = 0; c < nspec; c++)
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        size_t p = pidx<T>(c, iz, ix);
        const T f = a->q.w[ix] * a->q.w[iz] * a->jac[p];
        if (a->medium[c] == 1) {
          // include/medium/dim2/acoustic/isotropic/mass_matrix.tpp: 1 / kappa
          T m0 = static_cast<T>(1.0) / a->kappa[p];
          m0 *= f;
          if (a->bnd_tag[p])
            stacey_traction_ac<T>(a->bnd_nx[p], a->bnd_nz[p], a->bnd_w[p], a->rho_inv[p], a->kappa[p],
                                  velocity_factor, m0);
          mass_ac[a->index_ac[a->ibool[p]]] += m0;
          continue;
        }
        T m0 = a->rho[p], m1 = a->rho[p];
        m0 *= f;
        m1 *= f;
        if (a->bnd_tag[p])
          stacey_traction<T>(a->bnd_nx[p], a->bnd_nz[p], a->bnd_w[p], a->kappa[p], a->mu[p],
                             a->rho[p], velocity_factor, velocity_factor, m0, m1);
        const int g = a->medium_index[a->ibool[p]];
        mass[g] += m0;
        mass[static_cast<size_t>(nglob) + g] += m1;
      }
  a->mass_inverse.resize(2 * static_cast<size_t>(nglob));
  for (size_t i = 0; i < mass.size(); i++) a->mass_inverse[i] = static_cast<T>(1.0) / mass[i];
  a->mass_inverse_ac.resize(mass_ac.size());
  for (size_t i = 0; i < mass_ac.size(); i++) a->mass_inverse_ac[i] = static_cast<T>(1.0) / mass_ac[i];
}

// ---------------------------------------------------------------------------
// locate_point: src/algorithms/locate_point.cpp:8-176
// ---------------------------------------------------------------------------
template <class T> struct Local {
  int ispec;
  T xi, gamma;
};

template <class T> Local<T> locate_point(const Assembly<T> &a, T gx, T gz) {
  const int nspec = a.nspec, ngnod = a.ngnod;
  // rough_location (locate_point.cpp:8-40)
  T dist_min = std::numeric_limits<T>::max();
  int ispec_g = 0, ix_g = 0, iz_g = 0;
  for (int ispec = 0; ispec < nspec; ispec++)
    for (int j = 0; j < NGLL; j++)
      for (int i = 0; i < NGLL; i++) {
        size_t p = pidx<T>(ispec, j, i);
        const T dx = gx - a.coord[2 * p], dz = gz - a.coord[2 * p + 1];
        const T d = std::sqrt(dx * dx + dz * dz);
        if (d < dist_min) {
          ispec_g = ispec;
          ix_g = i;
          iz_g = j;
          dist_min = d;
        }
      }
  // get_best_candidates (locate_point.cpp:42-77)
  auto ib = [&](int e, int j, int i) { return a.ibool[pidx<T>(e, j, i)]; };
  std::vector<int> iglob_guess = {ib(ispec_g, 0, 0), ib(ispec_g, NGLL - 1, 0), ib(ispec_g, 0, NGLL - 1),
                                  ib(ispec_g, NGLL - 1, NGLL - 1)};
  std::vector<int> cand = {ispec_g};
  for (int e = 0; e < nspec; e++) {
    if (e == ispec_g) continue;
    for (int j : {0, NGLL - 1})
      for (int i : {0, NGLL - 1})
        if (std::find(iglob_guess.begin(), iglob_guess.end(), ib(e, j, i)) != iglob_guess.end())
          if (!cand.empty() && cand.back() != e) cand.push_back(e);
  }
  // get_best_location + selection (locate_point.cpp:79-176)
  std::vector<T> sh(ngnod), d0(ngnod), d1(ngnod);
  auto locations = [&](int e, T xi, T gamma, T &x, T &z) {
    shape2d<T>(xi, gamma, ngnod, sh.data());
    x = 0.0;
    z = 0.0;
    for (int in = 0; in < ngnod; in++) {
      x += sh[in] * a.ctrl[(static_cast<size_t>(e) * ngnod + in) * 2];
      z += sh[in] * a.ctrl[(static_cast<size_t>(e) * ngnod + in) * 2 + 1];
    }
  };
  T final_dist = std::numeric_limits<T>::max();
  Local<T> out{0, 0, 0};
  for (int e : cand) {
    T xi = a.q.xi[ix_g], gamma = a.q.xi[iz_g];
    for (int it = 0; it < 100; it++) {
      T x, z;
      locations(e, xi, gamma, x, z);
      dshape2d<T>(xi, gamma, ngnod, d0.data(), d1.data());
      T xxi = 0.0, zxi = 0.0, xgamma = 0.0, zgamma = 0.0;
      for (int in = 0; in < ngnod; in++) {
        const T cx = a.ctrl[(static_cast<size_t>(e) * ngnod + in) * 2];
        const T cz = a.ctrl[(static_cast<size_t>(e) * ngnod + in) * 2 + 1];
        xxi += d0[in] * cx;
        zxi += d0[in] * cz;
        xgamma += d1[in] * cx;
        zgamma += d1[in] * cz;
      }
      const T jacobian = xxi * zgamma - xgamma * zxi;
      const T xix = zgamma / jacobian, gammax = -zxi / jacobian;
      const T xiz = -xgamma / jacobian, gammaz = xxi / jacobian;
      const T dx = -(x - gx), dz = -(z - gz);
      xi += xix * dx + xiz * dz;
      gamma += gammax * dx + gammaz * dz;
      if (xi > 1.01) xi = 1.01;
      if (xi < -1.01) xi = -1.01;
      if (gamma > 1.01) gamma = 1.01;
      if (gamma < -1.01) gamma = -1.01;
    }
    T x, z;
    locations(e, xi, gamma, x, z);
    const T dx = gx - x, dz = gz - z;
    const T dist = std::sqrt(dx * dx + dz * dz);
    if (dist < final_dist) {
      out = {e, xi, gamma};
      final_dist = dist;
    }
    ix_g = static_cast<int>(NGLL / 2.0);
    iz_g = static_cast<int>(NGLL / 2.0);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Time loop: include/solver/time_marching.tpp:9-124 (forward, elastic only),
// Newmark include/timescheme/newmark.tpp:11-116, kernels
// include/kokkos_kernels/domain_kernels.hpp:69-126.
// ---------------------------------------------------------------------------
struct SourceIn {
  int type;     // 0 force, 1 moment tensor
  double x, z;
  double angle;                 // force
  double mxx, mzz, mxz;         // moment tensor
  double f0, tshift, factor;    // Ricker
  int stf = 0;                  // 0 Ricker, 1 Dirac
};

struct ReceiverIn {
  double x, z, angle;
};

// Ricker, src/source_time_function/ricker.cpp:39-50 and utilities.cpp:27-32.
template <class T> T ricker(T t, T f0, T factor) {
  const T pi = static_cast<T>(M_PI);
  const T a = pi * pi * f0 * f0;
  const T d2 = (1.0 - 2.0 * a * t * t) * std::exp(-a * t * t);
  return -1.0 * factor * d2;
}

// Dirac, src/source_time_function/dirac.cpp:24-48 and utilities.cpp:6-12:
// -factor times a Gaussian of half duration 1/f0, f0 = 1 / (10 dt).
template <class T> T dirac(T t, T f0, T factor) {
  const T pi = static_cast<T>(M_PI);
  const T a = pi * pi * f0 * f0;
  const T g = -1.0 * std::exp(-a * t * t) / (2.0 * a);
  return -1.0 * factor * g;
}
template <class T> T dirac_f0(T dt) { return static_cast<T>(1.0 / (10.0 * dt)); }

// One element's weak-form internal force (stiffness kernel body,
// include/kokkos_kernels/impl/compute_stiffness_interaction.tpp:143-259):
// gradient (include/algorithms/gradient.hpp:81-103), stress
// (include/medium/dim2/elastic/isotropic/stress.hpp:25-48), F = J sigma d
// (core/specfem/point/stress.hpp:118-139), divergence
// (include/algorithms/divergence.hpp:86-98), negation, Stacey.
template <class T>
void element_force(const Assembly<T> &a, int e, const T *ux, const T *uz, const T *vx_g, const T *vz_g,
                   T out[NGLL][NGLL][2]) {
  const auto &q = a.q;
  T lx[NGLL][NGLL], lz[NGLL][NGLL];
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) {
      const int g = a.medium_index[a.ibool[pidx<T>(e, iz, ix)]];
      lx[iz][ix] = ux[g];
      lz[iz][ix] = uz[g];
    }
  T F[NGLL][NGLL][2][2];
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) {
      T dxi[2] = {0.0, 0.0}, dgam[2] = {0.0, 0.0};
      for (int l = 0; l < NGLL; l++) {
        dxi[0] += q.hprime[ix][l] * lx[iz][l];
        dgam[0] += q.hprime[iz][l] * lx[l][ix];
        dxi[1] += q.hprime[ix][l] * lz[iz][l];
        dgam[1] += q.hprime[iz][l] * lz[l][ix];
      }
      const size_t p = pidx<T>(e, iz, ix);
      const T xix = a.xix[p], xiz = a.xiz[p], gammax = a.gammax[p], gammaz = a.gammaz[p], J = a.jac[p];
      T du[2][2];
      for (int c = 0; c < 2; c++) {
        du[c][0] = xix * dxi[c] + gammax * dgam[c];
        du[c][1] = xiz * dxi[c] + gammaz * dgam[c];
      }
      const T mu = a.mu[p];
      const T lp2m = a.kappa[p] + static_cast<T>(4.0 / 3.0) * mu;
      const T lambda = lp2m - static_cast<T>(2.0) * mu;
      const T sxx = lp2m * du[0][0] + lambda * du[1][1];
      const T szz = lp2m * du[1][1] + lambda * du[0][0];
      const T sxz = mu * (du[0][1] + du[1][0]);
      const T Tm[2][2] = {{sxx, sxz}, {sxz, szz}};
      for (int c = 0; c < 2; c++) {
        F[iz][ix][c][0] = J * (Tm[c][0] * xix + Tm[c][1] * xiz);
        F[iz][ix][c][1] = J * (Tm[c][0] * gammax + Tm[c][1] * gammaz);
      }
    }
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) {
      T t1[2] = {0.0, 0.0}, t2[2] = {0.0, 0.0};
      for (int l = 0; l < NGLL; l++) {
        for (int c = 0; c < 2; c++) t1[c] += F[iz][l][c][0] * q.hprime_wgll[ix][l];
        for (int c = 0; c < 2; c++) t2[c] += F[l][ix][c][1] * q.hprime_wgll[iz][l];
      }
      T r[2];
      for (int c = 0; c < 2; c++) r[c] = q.w[iz] * t1[c] + q.w[ix] * t2[c];
      for (int c = 0; c < 2; c++) r[c] *= static_cast<T>(-1.0);
      const size_t p = pidx<T>(e, iz, ix);
      if (a.bnd_tag[p]) {
        const int g = a.medium_index[a.ibool[p]];
        stacey_traction<T>(a.bnd_nx[p], a.bnd_nz[p], a.bnd_w[p], a.kappa[p], a.mu[p], a.rho[p], vx_g[g],
                           vz_g[g], r[0], r[1]);
      }
      out[iz][ix][0] = r[0];
      out[iz][ix][1] = r[1];
    }
}

// Point forcing of one wavefield: source arrays (ncomp, 5, 5), time functions
// stf(istep, comp) (include/compute/sources/source_medium.tpp:19-38) and the
// application order (element type, none then Stacey, then source order;
// src/compute/compute_sources.cpp:232-270).
template <class T> struct Forcing {
  int nstep = 0;
  std::vector<int> ispec;
  std::vector<T> arr;  // (nsrc, 2, 5, 5)
  std::vector<T> stf;  // (nsrc, nstep, 2)
  std::vector<int> order;
  // Element type groups of the source kernel launches (domain_kernels.hpp:
  // 81-95): boundary tags none, free surface, Stacey, composite; a medium's
  // update applies its own sources only.
  void finish_order(const Assembly<T> &a) {
    order.clear();
    for (int pass = 0; pass < 4; pass++)
      for (size_t s = 0; s < ispec.size(); s++)
        if (a.btag[ispec[s]] == pass) order.push_back(static_cast<int>(s));
  }
};

// Forward sources.  t0 = min over sources of (-1.2/f0 + tshift), then each
// source's tshift becomes its t0 - min_t0 (src/io/sources.cpp:128-158,
// ricker.cpp:15-20, ricker.hpp:49).
// A user start time (Assembly::start_time != 0) is used as t0 instead, and
// the sources keep their own tshift; it must not be later than
// min t0 - min tshift (src/io/sources.cpp:141-155).
template <class T>
Forcing<T> make_sources(const Assembly<T> &a, int nsrc, const SourceIn *src, int nstep, T dt, T *t0_out) {
  T min_t0 = std::numeric_limits<T>::max(), min_tshift = std::numeric_limits<T>::max();
  std::vector<T> src_t0(nsrc);
  for (int s = 0; s < nsrc; s++) {
    const T f0s = src[s].stf == 1 ? dirac_f0<T>(dt) : static_cast<T>(src[s].f0);
    T hdur = 1.0 / f0s;
    src_t0[s] = -1.2 * hdur + static_cast<T>(src[s].tshift);
    if (src_t0[s] < min_t0) min_t0 = src_t0[s];
    if (static_cast<T>(src[s].tshift) < min_tshift) min_tshift = static_cast<T>(src[s].tshift);
  }
  const T user = static_cast<T>(a.start_time);
  const bool user_defined = std::abs(user) > std::numeric_limits<T>::epsilon();
  if (user_defined && user > min_t0 - min_tshift)
    throw std::invalid_argument("start time later than the earliest source onset allows");
  const T t0 = user_defined ? user : min_t0;
  if (t0_out) *t0_out = t0;
  Forcing<T> f;
  f.nstep = nstep;
  f.ispec.resize(nsrc);
  f.arr.assign(static_cast<size_t>(nsrc) * 2 * NGLL * NGLL, 0);
  f.stf.assign(static_cast<size_t>(nsrc) * nstep * 2, 0);
  for (int s = 0; s < nsrc; s++) {
    auto loc = locate_point<T>(a, static_cast<T>(src[s].x), static_cast<T>(src[s].z));
    f.ispec[s] = loc.ispec;
    T *arr = &f.arr[static_cast<size_t>(s) * 2 * NGLL * NGLL];
    T hx[NGLL], hpx[NGLL], hg[NGLL], hpg[NGLL];
    lagrange_interpolants<T>(loc.xi, a.q.xi, hx, hpx);
    lagrange_interpolants<T>(loc.gamma, a.q.xi, hg, hpg);
    if (a.medium[loc.ispec] == 1) {
      // force source in an acoustic element: the Lagrange product alone
      // (src/source/force_source.cpp:51-56); moment tensors act on elastic media only
      if (src[s].type != 0) throw std::invalid_argument("moment-tensor source in an acoustic element");
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) arr[iz * NGLL + ix] = hx[ix] * hg[iz];
    } else if (src[s].type == 0) {
      // force source, src/source/force_source.cpp:70-78
      const T ang = static_cast<T>(src[s].angle);
      const T pi = static_cast<T>(M_PI);
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const T hl = hx[ix] * hg[iz];
          arr[iz * NGLL + ix] = std::sin(pi / 180 * ang) * hl;
          arr[NGLL * NGLL + iz * NGLL + ix] = -1.0 * std::cos(pi / 180 * ang) * hl;
        }
    } else {
      // moment tensor source, src/source/moment_tensor_source.cpp:88-130:
      // the GLL-point derivatives are interpolated to the source location
      // (include/algorithms/interpolate.hpp), then M contracts the gradient of
      // the Lagrange product.
      const int e = loc.ispec;
      T xixs = 0, xizs = 0, gammaxs = 0, gammazs = 0;
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const T hl = hx[ix] * hg[iz];
          const size_t p = pidx<T>(e, iz, ix);
          xixs += hl * a.xix[p];
          gammaxs += hl * a.gammax[p];
          xizs += hl * a.xiz[p];
          gammazs += hl * a.gammaz[p];
        }
      const T Mxx = static_cast<T>(src[s].mxx), Mzz = static_cast<T>(src[s].mzz), Mxz = static_cast<T>(src[s].mxz);
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const T dsx = (hpx[ix] * xixs) * hg[iz] + hx[ix] * (hpg[iz] * gammaxs);
          const T dsz = (hpx[ix] * xizs) * hg[iz] + hx[ix] * (hpg[iz] * gammazs);
          arr[iz * NGLL + ix] = Mxx * dsx + Mxz * dsz;
          arr[NGLL * NGLL + iz * NGLL + ix] = Mxz * dsx + Mzz * dsz;
        }
    }
    const T tshift = user_defined ? static_cast<T>(src[s].tshift) : src_t0[s] - min_t0;
    const T f0s = src[s].stf == 1 ? dirac_f0<T>(dt) : static_cast<T>(src[s].f0);
    for (int i = 0; i < nstep; i++) {
      const T v = src[s].stf == 1 ? dirac<T>(t0 + i * dt - tshift, f0s, static_cast<T>(src[s].factor))
                                  : ricker<T>(t0 + i * dt - tshift, f0s, static_cast<T>(src[s].factor));
      f.stf[(static_cast<size_t>(s) * nstep + i) * 2] = v;
      f.stf[(static_cast<size_t>(s) * nstep + i) * 2 + 1] = v;
    }
  }
  f.finish_order(a);
  return f;
}

// Adjoint sources (src/source/adjoint_source.cpp:9-90): the Lagrange product
// at the location for both components, with an external time function per
// component (src/source_time_function/external.cpp) given as stf
// (nadj, nstep, 2) in forward time order.
template <class T>
Forcing<T> make_adjoint_sources(const Assembly<T> &a, int nadj, const ReceiverIn *pos, int nstep, const double *stf) {
  Forcing<T> f;
  f.nstep = nstep;
  f.ispec.resize(nadj);
  f.arr.assign(static_cast<size_t>(nadj) * 2 * NGLL * NGLL, 0);
  f.stf.resize(static_cast<size_t>(nadj) * nstep * 2);
  for (int s = 0; s < nadj; s++) {
    auto loc = locate_point<T>(a, static_cast<T>(pos[s].x), static_cast<T>(pos[s].z));
    f.ispec[s] = loc.ispec;
    T hx[NGLL], hg[NGLL];
    lagrange_interpolants<T>(loc.xi, a.q.xi, hx, nullptr);
    lagrange_interpolants<T>(loc.gamma, a.q.xi, hg, nullptr);
    T *arr = &f.arr[static_cast<size_t>(s) * 2 * NGLL * NGLL];
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        const T hl = hx[ix] * hg[iz];
        arr[iz * NGLL + ix] = hl;
        arr[NGLL * NGLL + iz * NGLL + ix] = hl;
      }
  }
  for (size_t i = 0; i < f.stf.size(); i++) f.stf[i] = static_cast<T>(stf[i]);
  f.finish_order(a);
  return f;
}

// Receivers: element and Lagrange products (src/compute/compute_receivers.cpp:59-83).
template <class T> struct Receivers {
  std::vector<int> ispec;
  std::vector<T> h;  // (nrec, 5, 5)
};
template <class T> Receivers<T> make_receivers(const Assembly<T> &a, int nrec, const ReceiverIn *rec) {
  Receivers<T> R;
  R.ispec.resize(nrec);
  R.h.resize(static_cast<size_t>(nrec) * NGLL * NGLL);
  for (int r = 0; r < nrec; r++) {
    auto loc = locate_point<T>(a, static_cast<T>(rec[r].x), static_cast<T>(rec[r].z));
    R.ispec[r] = loc.ispec;
    T hx[NGLL], hg[NGLL];
    lagrange_interpolants<T>(loc.xi, a.q.xi, hx, nullptr);
    lagrange_interpolants<T>(loc.gamma, a.q.xi, hg, nullptr);
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) R.h[(static_cast<size_t>(r) * NGLL + iz) * NGLL + ix] = hx[ix] * hg[iz];
  }
  return R;
}

// One simulation field (include/compute/fields/impl/field_impl.hpp:30-38):
// (comp, nglob) on the per-medium index.
template <class T> struct Field {
  int nglob = 0, ncomp = 2;
  std::vector<T> u, v, acc;
  explicit Field(int n, int nc = 2) : nglob(n), ncomp(nc), u(nc * static_cast<size_t>(n), 0), v(u), acc(u) {}
};

// Newmark predictor (newmark.tpp:60-116).  The backward field uses
// (-dt, -dt/2, dt^2/2) (newmark.tpp:225-242).
template <class T> void predictor(Field<T> &f, T deltat, T deltatover2, T deltasquareover2, bool serial) {
  const size_t n = f.ncomp * static_cast<size_t>(f.nglob);
#pragma omp parallel for if (!serial) schedule(static)
  for (size_t i = 0; i < n; i++) {
    T add_u = 0;
    add_u += deltat * f.v[i] + deltasquareover2 * f.acc[i];
    T add_v = 0;
    add_v += deltatover2 * f.acc[i];
    f.u[i] += add_u;
    f.v[i] += add_v;
    f.acc[i] = 0;
  }
}

// Mass division (divide_mass_matrix.tpp:42-49) and corrector (newmark.tpp:11-56;
// -dt/2 for the backward field, :204-222).
template <class T> void divide_correct(const Assembly<T> &a, Field<T> &f, T deltatover2, bool serial) {
  const size_t n = f.ncomp * static_cast<size_t>(f.nglob);
  const T *minv = f.ncomp == 1 ? a.mass_inverse_ac.data() : a.mass_inverse.data();
#pragma omp parallel for if (!serial) schedule(static)
  for (size_t i = 0; i < n; i++) {
    f.acc[i] = f.acc[i] * minv[i];
    T add_v = 0;
    add_v += deltatover2 * f.acc[i];
    f.v[i] += add_v;
  }
}

// Stacey-element record: the whole local contribution (stiffness + traction)
// of every point of every Stacey element, per step
// (compute_stiffness_interaction.tpp:251-255, boundary_medium_container.hpp:26-82),
// stored here as (nstep, Stacey element in compute order, 5, 5, 2).
enum class StaceyMode { normal, record, replay };

// Sources then stiffness (domain_kernels.hpp:69-126), elements in compute
// order (none, then Stacey).  record: also store the Stacey elements' local
// contributions at rec_step; replay: add rec_step's values for the Stacey
// elements instead of computing them (the backward field,
// compute_stiffness_interaction.tpp:119-136).
template <class T>
void update_forces(const Assembly<T> &a, Field<T> &f, const Forcing<T> &F, int istep, StaceyMode mode, T *rec_step,
                   bool serial) {
  const int nglob = f.nglob;
  T *ax = f.acc.data(), *az = f.acc.data() + nglob;
  const T *ux = f.u.data(), *uz = f.u.data() + nglob, *vx = f.v.data(), *vz = f.v.data() + nglob;
  for (int s : F.order) {
    if (a.medium[F.ispec[s]] != 0) continue;
    const T *arr = &F.arr[static_cast<size_t>(s) * 2 * NGLL * NGLL];
    const T sx = F.stf[(static_cast<size_t>(s) * F.nstep + istep) * 2];
    const T sz = F.stf[(static_cast<size_t>(s) * F.nstep + istep) * 2 + 1];
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        const int g = a.medium_index[a.ibool[pidx<T>(F.ispec[s], iz, ix)]];
        ax[g] += sx * arr[iz * NGLL + ix];
        az[g] += sz * arr[NGLL * NGLL + iz * NGLL + ix];
      }
  }
  auto element = [&](int e, T out[NGLL][NGLL][2]) {
    const bool st = a.is_stacey[e];
    T *rec = st && mode != StaceyMode::normal ? rec_step + static_cast<size_t>(e - a.nspec_none) * NGLL * NGLL * 2
                                               : nullptr;
    if (rec && mode == StaceyMode::replay) {
      for (int p = 0; p < NGLL * NGLL; p++) {
        out[p / NGLL][p % NGLL][0] = rec[2 * p];
        out[p / NGLL][p % NGLL][1] = rec[2 * p + 1];
      }
      return;
    }
    element_force<T>(a, e, ux, uz, vx, vz, out);
    if (rec)
      for (int p = 0; p < NGLL * NGLL; p++) {
        rec[2 * p] = out[p / NGLL][p % NGLL][0];
        rec[2 * p + 1] = out[p / NGLL][p % NGLL][1];
      }
  };
  if (serial) {
    T out[NGLL][NGLL][2];
    for (int e = 0; e < a.nspec; e++) {
      if (a.medium[e] != 0) continue;
      element(e, out);
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const int g = a.medium_index[a.ibool[pidx<T>(e, iz, ix)]];
          ax[g] += out[iz][ix][0];
          az[g] += out[iz][ix][1];
        }
    }
  } else {
#pragma omp parallel for schedule(static)
    for (int e = 0; e < a.nspec; e++) {
      if (a.medium[e] != 0) continue;
      T out[NGLL][NGLL][2];
      element(e, out);
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const int g = a.medium_index[a.ibool[pidx<T>(e, iz, ix)]];
#pragma omp atomic
          ax[g] += out[iz][ix][0];
#pragma omp atomic
          az[g] += out[iz][ix][1];
        }
    }
  }
}

// Pressure at the points of element e (include/medium/dim2/elastic/isotropic/
// wavefield.hpp:49-85): -(lambda + 2/3 mu) (du_x/dx + du_z/dz), the factor in
// double as the reference's double constant promotes it.
template <class T> void element_pressure(const Assembly<T> &a, const Field<T> &f, int e, T out[NGLL][NGLL]) {
  const int nglob = f.nglob;
  const auto &q = a.q;
  T lx[NGLL][NGLL], lz[NGLL][NGLL];
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) {
      const int g = a.medium_index[a.ibool[pidx<T>(e, iz, ix)]];
      lx[iz][ix] = f.u[g];
      lz[iz][ix] = f.u[static_cast<size_t>(nglob) + g];
    }
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) {
      T dxi[2] = {0.0, 0.0}, dgam[2] = {0.0, 0.0};
      for (int l = 0; l < NGLL; l++) {
        dxi[0] += q.hprime[ix][l] * lx[iz][l];
        dgam[0] += q.hprime[iz][l] * lx[l][ix];
        dxi[1] += q.hprime[ix][l] * lz[iz][l];
        dgam[1] += q.hprime[iz][l] * lz[l][ix];
      }
      const size_t p = pidx<T>(e, iz, ix);
      const T du00 = a.xix[p] * dxi[0] + a.gammax[p] * dgam[0];
      const T du11 = a.xiz[p] * dxi[1] + a.gammaz[p] * dgam[1];
      const T mu = a.mu[p];
      const T lp2m = a.kappa[p] + static_cast<T>(4.0 / 3.0) * mu;
      const T lambda = lp2m - static_cast<T>(2.0) * mu;
      out[iz][ix] = static_cast<T>(-1.0 * ((lambda + (2.0 / 3.0) * mu) * (du00 + du11)));
    }
}

// Acoustic receiver values at the points of element e
// (include/medium/dim2/acoustic/isotropic/wavefield.hpp:14-84): for
// displacement, velocity and acceleration the gradient of the potential (or
// its time derivatives) times 1/rho, for pressure -chi_dot_dot.
template <class T> void element_acoustic_values(const Assembly<T> &a, const Field<T> &ac, int e, int type,
                                                T out[NGLL][NGLL][2]) {
  const auto &q = a.q;
  const std::vector<T> &fld = type == 0 ? ac.u : (type == 1 ? ac.v : ac.acc);
  T l[NGLL][NGLL];
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) l[iz][ix] = fld[a.index_ac[a.ibool[pidx<T>(e, iz, ix)]]];
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) {
      if (type == 3) {
        out[iz][ix][0] = static_cast<T>(-1.0 * l[iz][ix]);
        out[iz][ix][1] = 0;
        continue;
      }
      T dxi = 0.0, dgam = 0.0;
      for (int k = 0; k < NGLL; k++) {
        dxi += q.hprime[ix][k] * l[iz][k];
        dgam += q.hprime[iz][k] * l[k][ix];
      }
      const size_t p = pidx<T>(e, iz, ix);
      const T du0 = a.xix[p] * dxi + a.gammax[p] * dgam;
      const T du1 = a.xiz[p] * dxi + a.gammaz[p] * dgam;
      out[iz][ix][0] = a.rho_inv[p] * du0;
      out[iz][ix][1] = a.rho_inv[p] * du1;
    }
}

// Seismograms (compute_seismogram.tpp:103-135; include/algorithms/interpolate.hpp:176-237).
// ac: the acoustic field of a coupled mesh (receivers in acoustic elements).
template <class T>
void seismograms(const Assembly<T> &a, const Field<T> &f, const Receivers<T> &R, int ntypes, const int *types,
                 int isig, double *seis_out, const Field<T> *ac = nullptr) {
  const int nglob = f.nglob, nrec = static_cast<int>(R.ispec.size());
  for (int it = 0; it < ntypes; it++) {
    for (int r = 0; r < nrec; r++) {
      if (a.medium[R.ispec[r]] != 1) continue;
      T w[NGLL][NGLL][2];
      element_acoustic_values<T>(a, *ac, R.ispec[r], types[it], w);
      T s0 = 0, s1 = 0;
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const T h = R.h[(static_cast<size_t>(r) * NGLL + iz) * NGLL + ix];
          s0 += h * w[iz][ix][0];
          s1 += h * w[iz][ix][1];
        }
      const size_t o = ((static_cast<size_t>(isig) * ntypes + it) * nrec + r) * 2;
      seis_out[o] = s0;
      seis_out[o + 1] = types[it] == 3 ? 0.0 : s1;
    }
    if (types[it] == 3) {
      for (int r = 0; r < nrec; r++) {
        if (a.medium[R.ispec[r]] != 0) continue;
        T pr[NGLL][NGLL];
        element_pressure<T>(a, f, R.ispec[r], pr);
        T s0 = 0;
        for (int iz = 0; iz < NGLL; iz++)
          for (int ix = 0; ix < NGLL; ix++) s0 += R.h[(static_cast<size_t>(r) * NGLL + iz) * NGLL + ix] * pr[iz][ix];
        const size_t o = ((static_cast<size_t>(isig) * ntypes + it) * nrec + r) * 2;
        seis_out[o] = s0;
        seis_out[o + 1] = 0.0;
      }
      continue;
    }
    const std::vector<T> &fld = types[it] == 0 ? f.u : (types[it] == 1 ? f.v : f.acc);
    const T *fx = fld.data(), *fz = fld.data() + nglob;
    for (int r = 0; r < nrec; r++) {
      if (a.medium[R.ispec[r]] != 0) continue;
      T s0 = 0, s1 = 0;
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const int g = a.medium_index[a.ibool[pidx<T>(R.ispec[r], iz, ix)]];
          const T h = R.h[(static_cast<size_t>(r) * NGLL + iz) * NGLL + ix];
          s0 += h * fx[g];
          s1 += h * fz[g];
        }
      const size_t o = ((static_cast<size_t>(isig) * ntypes + it) * nrec + r) * 2;
      seis_out[o] = s0;
      seis_out[o + 1] = s1;
    }
  }
}

// Snapshot of one field, (3, 2, nglob): u, v, a.
template <class T> void store_state(const Field<T> &f, T *out) {
  const size_t n = 2 * static_cast<size_t>(f.nglob);
  std::copy(f.u.begin(), f.u.end(), out);
  std::copy(f.v.begin(), f.v.end(), out + n);
  std::copy(f.acc.begin(), f.acc.end(), out + 2 * n);
}
template <class T> void load_state(Field<T> &f, const T *in) {
  const size_t n = 2 * static_cast<size_t>(f.nglob);
  std::copy(in, in + n, f.u.begin());
  std::copy(in + n, in + 2 * n, f.v.begin());
  std::copy(in + 2 * n, in + 3 * n, f.acc.begin());
}

// Forward simulation (time_marching.tpp:9-124).  Seismogram output layout:
// (isig, itype, irec, 2) as the reference's (max_sig_step, ntypes, nrec, 2)
// array before rotation; types 0 displacement, 1 velocity, 2 acceleration.
// record: optional Stacey-element record (nstep, nstacey, 5, 5, 2).
template <class T>
void run_forward(const Assembly<T> &a, int nsrc, const SourceIn *src, int nrec, const ReceiverIn *rec, int nstep,
                 double dt_in, int nsb, int ntypes, const int *types, double *seis_out, double *t0_out,
                 int nthreads, T *final_fields /* optional (3, 2, nglob) */, T *record = nullptr,
                 int snapshot_every = 0, T *snapshots = nullptr /* (nsnap, 3, 2, nglob) */) {
  const T dt = static_cast<T>(dt_in);
  const T deltat = dt, deltatover2 = dt / 2.0, deltasquareover2 = dt * dt / 2.0;
  T t0 = 0;
  const Forcing<T> S = make_sources<T>(a, nsrc, src, nstep, dt, &t0);
  if (t0_out) *t0_out = t0;
  const Receivers<T> R = make_receivers<T>(a, nrec, rec);
  Field<T> f(a.nglob);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const bool serial = nthreads <= 1;
  const size_t rec_stride = static_cast<size_t>(a.nspec - a.nspec_none) * NGLL * NGLL * 2;
  int isig = 0;
  for (int istep = 0; istep < nstep; istep++) {
    predictor<T>(f, deltat, deltatover2, deltasquareover2, serial);
    update_forces<T>(a, f, S, istep, record ? StaceyMode::record : StaceyMode::normal,
                     record ? record + istep * rec_stride : nullptr, serial);
    divide_correct<T>(a, f, deltatover2, serial);
    if (istep % nsb == 0) seismograms<T>(a, f, R, ntypes, types, isig++, seis_out);
    if (snapshots && snapshot_every > 0 && istep % snapshot_every == 0 && istep < nstep - 1)
      store_state<T>(f, snapshots + static_cast<size_t>(istep / snapshot_every) * 6 * a.nglob);
  }
  if (final_fields) {
    std::copy(f.u.begin(), f.u.end(), final_fields);
    std::copy(f.v.begin(), f.v.end(), final_fields + 2 * static_cast<size_t>(a.nglob));
    std::copy(f.acc.begin(), f.acc.end(), final_fields + 4 * static_cast<size_t>(a.nglob));
  }
}

// ---------------------------------------------------------------------------
// Acoustic medium and elastic-acoustic coupling (2-D forward).
// ---------------------------------------------------------------------------

// Edge iterator of the coupling kernel (include/policies/edge.hpp, struct
// edge<dim2>): the self and coupled point of edge point ipoint, as (iz, ix).
inline void self_point(int edge, int ip, int &iz, int &ix) {
  const int n = NGLL - 1;
  switch (edge) {
    case BOTTOM: iz = 0; ix = ip; break;
    case TOP: iz = n; ix = n - ip; break;
    case LEFT: iz = ip; ix = 0; break;
    default: iz = n - ip; ix = n; break;  // RIGHT
  }
}
inline void coupled_point(int edge, int ip, int &iz, int &ix) {
  const int n = NGLL - 1;
  switch (edge) {
    case BOTTOM: iz = 0; ix = n - ip; break;
    case TOP: iz = n; ix = ip; break;
    case LEFT: iz = n - ip; ix = n; break;
    default: iz = ip; ix = 0; break;  // RIGHT
  }
}

// Acoustic element force (compute_stiffness_interaction.tpp:143-259 with
// the acoustic stress T = rho_inverse grad chi,
// include/medium/dim2/acoustic/isotropic/stress.hpp), Stacey
// (stacey.tpp:172-213) then the free surface (dirichlet.hpp: the point's
// whole contribution set to zero).
template <class T>
void element_force_ac(const Assembly<T> &a, int e, const Field<T> &ac, T out[NGLL][NGLL]) {
  const auto &q = a.q;
  T l[NGLL][NGLL];
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) l[iz][ix] = ac.u[a.index_ac[a.ibool[pidx<T>(e, iz, ix)]]];
  T F[NGLL][NGLL][2];
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) {
      T dxi = 0.0, dgam = 0.0;
      for (int k = 0; k < NGLL; k++) {
        dxi += q.hprime[ix][k] * l[iz][k];
        dgam += q.hprime[iz][k] * l[k][ix];
      }
      const size_t p = pidx<T>(e, iz, ix);
      const T xix = a.xix[p], xiz = a.xiz[p], gammax = a.gammax[p], gammaz = a.gammaz[p], J = a.jac[p];
      const T du0 = xix * dxi + gammax * dgam;
      const T du1 = xiz * dxi + gammaz * dgam;
      const T T0 = a.rho_inv[p] * du0, T1 = a.rho_inv[p] * du1;
      F[iz][ix][0] = J * (T0 * xix + T1 * xiz);
      F[iz][ix][1] = J * (T0 * gammax + T1 * gammaz);
    }
  for (int iz = 0; iz < NGLL; iz++)
    for (int ix = 0; ix < NGLL; ix++) {
      T t1 = 0.0, t2 = 0.0;
      for (int k = 0; k < NGLL; k++) {
        t1 += F[iz][k][0] * q.hprime_wgll[ix][k];
        t2 += F[k][ix][1] * q.hprime_wgll[iz][k];
      }
      T r = q.w[iz] * t1 + q.w[ix] * t2;
      r *= static_cast<T>(-1.0);
      const size_t p = pidx<T>(e, iz, ix);
      if (a.bnd_tag[p])
        stacey_traction_ac<T>(a.bnd_nx[p], a.bnd_nz[p], a.bnd_w[p], a.rho_inv[p], a.kappa[p],
                              ac.v[a.index_ac[a.ibool[p]]], r);
      if (a.dir_tag[p]) r = 0;
      out[iz][ix] = r;
    }
}

// update_wavefields<acoustic> (domain_kernels.hpp:69-126): coupling from the
// elastic displacement (coupled_interface.tpp: factor n.u, n the acoustic
// side's normal, i.e. minus the elastic one), sources (stf * lagrange /
// kappa, include/medium/dim2/acoustic/isotropic/source.hpp), stiffness.
template <class T>
void update_acoustic(const Assembly<T> &a, Field<T> &ac, const Field<T> &el, const Forcing<T> &F, int istep) {
  T *acc = ac.acc.data();
  const size_t ne = static_cast<size_t>(el.nglob);
  for (const auto &ce : a.coupling)
    for (int ip = 0; ip < NGLL; ip++) {
      int sz, sx, cz, cx;
      self_point(ce.edge_ac, ip, sz, sx);
      coupled_point(ce.edge_el, ip, cz, cx);
      const int gs = a.index_ac[a.ibool[pidx<T>(ce.e_ac, sz, sx)]];
      const int gc = a.medium_index[a.ibool[pidx<T>(ce.e_el, cz, cx)]];
      const T nx = -ce.nx[ip], nz = -ce.nz[ip];
      acc[gs] += ce.factor[ip] * (nx * el.u[gc] + nz * el.u[ne + gc]);
    }
  for (int s : F.order) {
    const int e = F.ispec[s];
    if (a.medium[e] != 1) continue;
    const T *arr = &F.arr[static_cast<size_t>(s) * 2 * NGLL * NGLL];
    const T sv = F.stf[(static_cast<size_t>(s) * F.nstep + istep) * 2];
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        const size_t p = pidx<T>(e, iz, ix);
        acc[a.index_ac[a.ibool[p]]] += sv * arr[iz * NGLL + ix] / a.kappa[p];
      }
  }
  T out[NGLL][NGLL];
  for (int e = 0; e < a.nspec; e++) {
    if (a.medium[e] != 1) continue;
    element_force_ac<T>(a, e, ac, out);
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) acc[a.index_ac[a.ibool[pidx<T>(e, iz, ix)]]] += out[iz][ix];
  }
}

// The elastic side of the coupling (coupled_interface.tpp): factor n chi_dot_dot.
template <class T> void couple_into_elastic(const Assembly<T> &a, Field<T> &el, const Field<T> &ac) {
  const size_t ne = static_cast<size_t>(el.nglob);
  for (const auto &ce : a.coupling)
    for (int ip = 0; ip < NGLL; ip++) {
      int sz, sx, cz, cx;
      self_point(ce.edge_el, ip, sz, sx);
      coupled_point(ce.edge_ac, ip, cz, cx);
      const int gs = a.medium_index[a.ibool[pidx<T>(ce.e_el, sz, sx)]];
      const int gc = a.index_ac[a.ibool[pidx<T>(ce.e_ac, cz, cx)]];
      el.acc[gs] += ce.factor[ip] * ce.nx[ip] * ac.acc[gc];
      el.acc[ne + gs] += ce.factor[ip] * ce.nz[ip] * ac.acc[gc];
    }
}

// Forward loop of a mesh with acoustic elements (time_marching.tpp:33-84):
// predictors, acoustic update and corrector, elastic update (coupling first)
// and corrector, seismograms.  Serial order throughout.
template <class T>
void run_forward_coupled(const Assembly<T> &a, int nsrc, const SourceIn *src, int nrec, const ReceiverIn *rec,
                         int nstep, double dt_in, int nsb, int ntypes, const int *types, double *seis_out,
                         double *t0_out, T *final_fields /* optional: (3, 2, nglob_el) then (3, nglob_ac) */) {
  const T dt = static_cast<T>(dt_in);
  const T deltat = dt, deltatover2 = dt / 2.0, deltasquareover2 = dt * dt / 2.0;
  T t0 = 0;
  const Forcing<T> S = make_sources<T>(a, nsrc, src, nstep, dt, &t0);
  if (t0_out) *t0_out = t0;
  const Receivers<T> R = make_receivers<T>(a, nrec, rec);
  Field<T> el(a.nglob_el), ac(a.nglob_ac, 1);
  int isig = 0;
  for (int istep = 0; istep < nstep; istep++) {
    predictor<T>(ac, deltat, deltatover2, deltasquareover2, true);
    predictor<T>(el, deltat, deltatover2, deltasquareover2, true);
    update_acoustic<T>(a, ac, el, S, istep);
    divide_correct<T>(a, ac, deltatover2, true);
    couple_into_elastic<T>(a, el, ac);
    update_forces<T>(a, el, S, istep, StaceyMode::normal, nullptr, true);
    divide_correct<T>(a, el, deltatover2, true);
    if (istep % nsb == 0) seismograms<T>(a, el, R, ntypes, types, isig++, seis_out, &ac);
  }
  if (final_fields) {
    const size_t ne = 2 * static_cast<size_t>(a.nglob_el), na = a.nglob_ac;
    std::copy(el.u.begin(), el.u.end(), final_fields);
    std::copy(el.v.begin(), el.v.end(), final_fields + ne);
    std::copy(el.acc.begin(), el.acc.end(), final_fields + 2 * ne);
    std::copy(ac.u.begin(), ac.u.end(), final_fields + 3 * ne);
    std::copy(ac.v.begin(), ac.v.end(), final_fields + 3 * ne + na);
    std::copy(ac.acc.begin(), ac.acc.end(), final_fields + 3 * ne + 2 * na);
  }
}

// Fréchet derivatives of one step (compute_material_derivatives.tpp:17-154,
// include/medium/dim2/elastic/isotropic/frechet_derivative.hpp:14-138):
// kernels (6, nspec, 5, 5) in compute order, rho, mu, kappa, rhop, alpha,
// beta, accumulated.
template <class T>
void frechet_step(const Assembly<T> &a, const Field<T> &adj, const Field<T> &bwd, T dt, T *K, bool serial) {
  const int nglob = a.nglob;
  const size_t npts = static_cast<size_t>(a.nspec) * NGLL * NGLL;
  const auto &q = a.q;
#pragma omp parallel for if (!serial) schedule(static)
  for (int e = 0; e < a.nspec; e++) {
    T ad[2][NGLL][NGLL], bw[2][NGLL][NGLL];
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        const int g = a.medium_index[a.ibool[pidx<T>(e, iz, ix)]];
        for (int c = 0; c < 2; c++) {
          ad[c][iz][ix] = adj.u[static_cast<size_t>(c) * nglob + g];
          bw[c][iz][ix] = bwd.u[static_cast<size_t>(c) * nglob + g];
        }
      }
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        const size_t p = pidx<T>(e, iz, ix);
        const T xix = a.xix[p], xiz = a.xiz[p], gammax = a.gammax[p], gammaz = a.gammaz[p];
        T df[2][2], dg[2][2];  // du(c, j) = d u_c / d x_j (gradient.hpp:81-103)
        for (int c = 0; c < 2; c++) {
          T fxi = 0, fga = 0, gxi = 0, gga = 0;
          for (int l = 0; l < NGLL; l++) {
            fxi += q.hprime[ix][l] * ad[c][iz][l];
            fga += q.hprime[iz][l] * ad[c][l][ix];
            gxi += q.hprime[ix][l] * bw[c][iz][l];
            gga += q.hprime[iz][l] * bw[c][l][ix];
          }
          df[c][0] = xix * fxi + gammax * fga;
          df[c][1] = xiz * fxi + gammaz * fga;
          dg[c][0] = xix * gxi + gammax * gga;
          dg[c][1] = xiz * gxi + gammaz * gga;
        }
        const T mu = a.mu[p], rho = a.rho[p];
        const T lp2m = a.kappa[p] + static_cast<T>(4.0 / 3.0) * mu;
        const T kappa = lp2m - mu;
        const T ad_dsxx = df[0][0];
        const T ad_dsxz = static_cast<T>(0.5) * (df[1][0] + df[0][1]);
        const T ad_dszz = df[1][1];
        const T b_dsxx = dg[0][0];
        const T b_dsxz = static_cast<T>(0.5) * (dg[1][0] + dg[0][1]);
        const T b_dszz = dg[1][1];
        T kappa_kl = (ad_dsxx + ad_dszz) * (b_dsxx + b_dszz);
        T mu_kl = (ad_dsxx * b_dsxx + ad_dszz * b_dszz + static_cast<T>(2.0) * ad_dsxz * b_dsxz -
                   static_cast<T>(1.0 / 3.0) * kappa_kl);
        const int g = a.medium_index[a.ibool[p]];
        T rho_kl = adj.acc[g] * bwd.u[g] + adj.acc[static_cast<size_t>(nglob) + g] * bwd.u[static_cast<size_t>(nglob) + g];
        kappa_kl = static_cast<T>(-1.0) * kappa * dt * kappa_kl;
        mu_kl = static_cast<T>(-2.0) * mu * dt * mu_kl;
        rho_kl = static_cast<T>(-1.0) * rho * dt * rho_kl;
        const T rhop_kl = rho_kl + kappa_kl + mu_kl;
        const T beta_kl = static_cast<T>(2.0) * (mu_kl - static_cast<T>(4.0 / 3.0) * mu / kappa * kappa_kl);
        const T alpha_kl = static_cast<T>(2.0) * (static_cast<T>(1.0) + static_cast<T>(4.0 / 3.0) * mu / kappa) * kappa_kl;
        K[0 * npts + p] += rho_kl;
        K[1 * npts + p] += mu_kl;
        K[2 * npts + p] += kappa_kl;
        K[3 * npts + p] += rhop_kl;
        K[4 * npts + p] += alpha_kl;
        K[5 * npts + p] += beta_kl;
      }
  }
}

// Combined adjoint + backward loop (time_marching.tpp:126-247): istep from
// nstep - 1 down to 0; the adjoint field steps forward in time with the
// adjoint sources, the backward field steps with -dt from the forward run's
// final state (copied in after the first backward step), replaying the
// recorded Stacey-element contributions.  Kernels accumulate every step;
// backward seismograms are sampled at istep % nsb == 0 (sample 0 is the
// latest time).
template <class T>
void run_combined(const Assembly<T> &a, int nsrc, const SourceIn *src, int nadj, const ReceiverIn *adj_pos,
                  const double *adj_stf, int nrec, const ReceiverIn *rec, int nstep, double dt_in, int nsb,
                  int ntypes, const int *types, const T *final_fields, const T *record, T *K, double *seis_out,
                  int nthreads, int snapshot_every = 0, const T *snapshots = nullptr) {
  const T dt = static_cast<T>(dt_in);
  const T deltat = dt, deltatover2 = dt / 2.0, deltasquareover2 = dt * dt / 2.0;
  const Forcing<T> S = make_sources<T>(a, nsrc, src, nstep, dt, nullptr);
  const Forcing<T> A = make_adjoint_sources<T>(a, nadj, adj_pos, nstep, adj_stf);
  const Receivers<T> R = make_receivers<T>(a, nrec, rec);
  Field<T> adj(a.nglob), bwd(a.nglob);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const bool serial = nthreads <= 1;
  const size_t rec_stride = static_cast<size_t>(a.nspec - a.nspec_none) * NGLL * NGLL * 2;
  std::vector<T> scratch;
  int isig = 0;
  for (int istep = nstep - 1; istep >= 0; istep--) {
    predictor<T>(adj, deltat, deltatover2, deltasquareover2, serial);
    update_forces<T>(a, adj, A, istep, StaceyMode::normal, nullptr, serial);
    divide_correct<T>(a, adj, deltatover2, serial);

    predictor<T>(bwd, -deltat, -deltatover2, deltasquareover2, serial);
    scratch.assign(record + istep * rec_stride, record + (istep + 1) * rec_stride);
    update_forces<T>(a, bwd, S, istep, StaceyMode::replay, scratch.data(), serial);
    divide_correct<T>(a, bwd, -deltatover2, serial);
    if (istep == nstep - 1) load_state<T>(bwd, final_fields);
    // checkpoint: the backward state after istep is the forward state after
    // step istep; reset it from the snapshot taken in the forward run
    if (snapshots && snapshot_every > 0 && istep % snapshot_every == 0 && istep < nstep - 1)
      load_state<T>(bwd, snapshots + static_cast<size_t>(istep / snapshot_every) * 6 * a.nglob);
    frechet_step<T>(a, adj, bwd, dt, K, serial);
    if (istep % nsb == 0) seismograms<T>(a, bwd, R, ntypes, types, isig++, seis_out);
  }
}

// Apply the assembled stiffness operator once: out = sum_e (-K_e u) with the
// Stacey term driven by v (used by the brute-force weak-form checks).
template <class T> void apply_stiffness(const Assembly<T> &a, const T *u, const T *v, T *out) {
  const int nglob = a.nglob;
  std::fill(out, out + 2 * static_cast<size_t>(nglob), static_cast<T>(0));
  T o[NGLL][NGLL][2];
  for (int e = 0; e < a.nspec; e++) {
    element_force<T>(a, e, u, u + nglob, v, v + nglob, o);
    for (int iz = 0; iz < NGLL; iz++)
      for (int ix = 0; ix < NGLL; ix++) {
        const int g = a.medium_index[a.ibool[pidx<T>(e, iz, ix)]];
        out[g] += o[iz][ix][0];
        out[static_cast<size_t>(nglob) + g] += o[iz][ix][1];
      }
  }
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C interface for ctypes (tests, smoke, bench cpu_baseline)
// ---------------------------------------------------------------------------
using namespace orc;

namespace {
thread_local std::string g_err;
template <class F> int guard(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}
struct Handle {
  int precision;  // 4 or 8
  void *a;
};
template <class T> Assembly<T> &asm_of(void *h) { return *static_cast<Assembly<T> *>(static_cast<Handle *>(h)->a); }
}  // namespace

#define DISPATCH(h, ...)                                       \
  do {                                                         \
    if (static_cast<Handle *>(h)->precision == 4) {            \
      using T = float;                                         \
      __VA_ARGS__;                                             \
    } else {                                                   \
      using T = double;                                        \
      __VA_ARGS__;                                             \
    }                                                          \
  } while (0)

extern "C" {

const char *orc_last_error() { return g_err.c_str(); }

int orc2d_create(int precision, int nspec, int ngnod, int npgeo, const double *coorg, const int *knods,
                 const int *kmato, int nmat, const double *rho, const double *cp, const double *cs, int nabs,
                 const int *abs_ispec, const int *abs_type, int nafs, const int *afs_ispec, const int *afs_type,
                 int nfs, const int *fs_acoustic, const int *fs_elastic, int chunk, double dt, void **out) {
  return guard([&] {
    MeshIn m{nspec, ngnod, npgeo, coorg, knods, kmato, nmat, rho, cp, cs, nabs, abs_ispec, abs_type,
             nafs, afs_ispec, afs_type, nfs, fs_acoustic, fs_elastic};
    auto *h = new Handle{precision, nullptr};
    if (precision == 4)
      h->a = build_assembly<float>(m, chunk, dt);
    else if (precision == 8)
      h->a = build_assembly<double>(m, chunk, dt);
    else
      throw std::invalid_argument("precision must be 4 or 8");
    *out = h;
  });
}

void orc2d_destroy(void *h) {
  if (!h) return;
  DISPATCH(h, delete &asm_of<T>(h));
  delete static_cast<Handle *>(h);
}

int orc2d_nglob(void *h) {
  int n = 0;
  DISPATCH(h, n = asm_of<T>(h).nglob);
  return n;
}

// (nglob of the elastic field, nglob of the acoustic field)
void orc2d_nglob_media(void *h, int *out2) {
  DISPATCH(h, {
    out2[0] = asm_of<T>(h).nglob_el;
    out2[1] = asm_of<T>(h).nglob_ac;
  });
}

// acoustic field index of every iglob (-1 off the acoustic elements), the
// medium of every compute element and the acoustic mass inverse (nglob_ac)
void orc2d_acoustic(void *h, int *index_ac, int *medium, double *mass_inverse_ac) {
  DISPATCH(h, {
    auto &a = asm_of<T>(h);
    if (index_ac) std::copy(a.index_ac.begin(), a.index_ac.end(), index_ac);
    if (medium) std::copy(a.medium.begin(), a.medium.end(), medium);
    if (mass_inverse_ac)
      for (size_t i = 0; i < a.mass_inverse_ac.size(); i++) mass_inverse_ac[i] = a.mass_inverse_ac[i];
  });
}

// ibool (nspec, 5, 5) compute order; medium index (nglob); compute_to_mesh (nspec)
void orc2d_numbering(void *h, int *ibool, int *medium_index, int *compute_to_mesh) {
  DISPATCH(h, {
    auto &a = asm_of<T>(h);
    if (ibool) std::copy(a.ibool.begin(), a.ibool.end(), ibool);
    if (medium_index) std::copy(a.medium_index.begin(), a.medium_index.end(), medium_index);
    if (compute_to_mesh) std::copy(a.compute_to_mesh.begin(), a.compute_to_mesh.end(), compute_to_mesh);
  });
}

// point coordinates (nspec,5,5,2), jacobian and derivatives (nspec,5,5)
void orc2d_geometry(void *h, double *coord, double *xix, double *xiz, double *gammax, double *gammaz,
                    double *jac) {
  DISPATCH(h, {
    auto &a = asm_of<T>(h);
    const size_t np = a.xix.size();
    for (size_t i = 0; i < np; i++) {
      if (coord) { coord[2 * i] = a.coord[2 * i]; coord[2 * i + 1] = a.coord[2 * i + 1]; }
      if (xix) xix[i] = a.xix[i];
      if (xiz) xiz[i] = a.xiz[i];
      if (gammax) gammax[i] = a.gammax[i];
      if (gammaz) gammaz[i] = a.gammaz[i];
      if (jac) jac[i] = a.jac[i];
    }
  });
}

// mass inverse (2, nglob) on the per-medium index
void orc2d_mass_inverse(void *h, double *out) {
  DISPATCH(h, {
    auto &a = asm_of<T>(h);
    for (size_t i = 0; i < a.mass_inverse.size(); i++) out[i] = a.mass_inverse[i];
  });
}

void orc2d_quadrature(void *h, double *xi, double *w, double *hprime) {
  DISPATCH(h, {
    auto &a = asm_of<T>(h);
    for (int i = 0; i < NGLL; i++) {
      xi[i] = a.q.xi[i];
      w[i] = a.q.w[i];
      for (int j = 0; j < NGLL; j++) hprime[i * NGLL + j] = a.q.hprime[i][j];
    }
  });
}

int orc2d_locate(void *h, double x, double z, int *ispec, double *xi, double *gamma) {
  return guard([&] {
    DISPATCH(h, {
      auto loc = locate_point<T>(asm_of<T>(h), static_cast<T>(x), static_cast<T>(z));
      *ispec = loc.ispec;
      *xi = loc.xi;
      *gamma = loc.gamma;
    });
  });
}

}  // extern "C"

namespace {
std::vector<SourceIn> sources_in(int nsrc, const double *src) {
  std::vector<SourceIn> S(nsrc);
  for (int s = 0; s < nsrc; s++) {
    const double *r = src + 11 * s;
    S[s] = SourceIn{static_cast<int>(r[0]), r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9],
                    static_cast<int>(r[10])};
  }
  return S;
}
std::vector<ReceiverIn> receivers_in(int nrec, const double *rec) {
  std::vector<ReceiverIn> R(nrec);
  for (int i = 0; i < nrec; i++) R[i] = ReceiverIn{rec[3 * i], rec[3 * i + 1], rec[3 * i + 2]};
  return R;
}
}  // namespace

extern "C" {

// sources: (nsrc, 11) rows [type, x, z, angle, mxx, mzz, mxz, f0, tshift, factor, stf (0 Ricker, 1 Dirac)]
// receivers: (nrec, 3) rows [x, z, angle]
int orc2d_forward(void *h, int nsrc, const double *src, int nrec, const double *rec, int nstep, double dt, int nsb,
                  int ntypes, const int *types, double *seis, double *t0, int nthreads, double *final_fields) {
  return guard([&] {
    auto S = sources_in(nsrc, src);
    auto R = receivers_in(nrec, rec);
    DISPATCH(h, {
      auto &a = asm_of<T>(h);
      std::vector<T> ff;
      if (a.coupled()) {
        // final fields: elastic (3, 2, nglob_el) then acoustic (3, nglob_ac)
        if (final_fields) ff.resize(6 * static_cast<size_t>(a.nglob_el) + 3 * static_cast<size_t>(a.nglob_ac));
        run_forward_coupled<T>(a, nsrc, S.data(), nrec, R.data(), nstep, dt, nsb, ntypes, types, seis, t0,
                               final_fields ? ff.data() : nullptr);
      } else {
        if (final_fields) ff.resize(6 * static_cast<size_t>(a.nglob));
        run_forward<T>(a, nsrc, S.data(), nrec, R.data(), nstep, dt, nsb, ntypes, types, seis, t0, nthreads,
                       final_fields ? ff.data() : nullptr);
      }
      if (final_fields)
        for (size_t i = 0; i < ff.size(); i++) final_fields[i] = ff[i];
    });
  });
}

// Sample given fields (3, 2, nglob) [u, v, a] at receivers rec (nrec, 3):
// out (ntypes, nrec, 2).
int orc2d_sample(void *h, int nrec, const double *rec, int ntypes, const int *types, const double *fields,
                 double *out) {
  return guard([&] {
    auto R = receivers_in(nrec, rec);
    DISPATCH(h, {
      auto &a = asm_of<T>(h);
      if (a.coupled()) throw std::invalid_argument("elastic meshes only");
      Field<T> f(a.nglob);
      const size_t n = 2 * static_cast<size_t>(a.nglob);
      for (size_t i = 0; i < n; i++) {
        f.u[i] = static_cast<T>(fields[i]);
        f.v[i] = static_cast<T>(fields[n + i]);
        f.acc[i] = static_cast<T>(fields[2 * n + i]);
      }
      const Receivers<T> RR = make_receivers<T>(a, nrec, R.data());
      seismograms<T>(a, f, RR, ntypes, types, 0, out);
    });
  });
}

// Replace the per-point properties (nspec, 5, 5) in mesh element order and
// rebuild the mass matrix (finite-difference checks of the kernels).
void orc2d_set_properties(void *h, const double *rho, const double *kappa, const double *mu) {
  DISPATCH(h, {
    auto &a = asm_of<T>(h);
    for (int c = 0; c < a.nspec; c++)
      for (int p = 0; p < NGLL * NGLL; p++) {
        const size_t src = static_cast<size_t>(a.compute_to_mesh[c]) * NGLL * NGLL + p;
        const size_t dst = static_cast<size_t>(c) * NGLL * NGLL + p;
        a.rho[dst] = static_cast<T>(rho[src]);
        a.kappa[dst] = static_cast<T>(kappa[src]);
        a.mu[dst] = static_cast<T>(mu[src]);
      }
    compute_mass(a);
  });
}

// Per-point properties and Jacobian weights w_i w_j J (nspec, 5, 5) in mesh
// element order.
void orc2d_properties(void *h, double *rho, double *kappa, double *mu, double *wj) {
  DISPATCH(h, {
    auto &a = asm_of<T>(h);
    for (int c = 0; c < a.nspec; c++)
      for (int iz = 0; iz < NGLL; iz++)
        for (int ix = 0; ix < NGLL; ix++) {
          const size_t dst = (static_cast<size_t>(a.compute_to_mesh[c]) * NGLL + iz) * NGLL + ix;
          const size_t src = pidx<T>(c, iz, ix);
          rho[dst] = a.rho[src];
          kappa[dst] = a.kappa[src];
          mu[dst] = a.mu[src];
          wj[dst] = a.q.w[ix] * a.q.w[iz] * a.jac[src];
        }
  });
}

void orc2d_set_start_time(void *h, double t) {
  DISPATCH(h, { asm_of<T>(h).start_time = t; });
}

int orc2d_nstacey(void *h) {
  int n = 0;
  DISPATCH(h, n = asm_of<T>(h).nspec - asm_of<T>(h).nspec_none);
  return n;
}


// Forward run that also returns the final fields (3, 2, nglob) and the
// Stacey-element record (nstep, nstacey, 5, 5, 2) for a combined run.
int orc2d_forward_record(void *h, int nsrc, const double *src, int nrec, const double *rec, int nstep, double dt,
                         int nsb, int ntypes, const int *types, double *seis, double *t0, int nthreads,
                         double *final_fields, double *record, int snapshot_every, double *snapshots) {
  return guard([&] {
    auto S = sources_in(nsrc, src);
    auto R = receivers_in(nrec, rec);
    DISPATCH(h, {
      auto &a = asm_of<T>(h);
      if (a.coupled()) throw std::invalid_argument("elastic meshes only");
      if (!a.is_stacey.empty())
        for (int e = 0; e < a.nspec; e++)
          if ((e >= a.nspec_none) != (a.is_stacey[e] != 0))
            throw std::logic_error("Stacey elements are not the last in compute order");
      std::vector<T> ff(6 * static_cast<size_t>(a.nglob));
      std::vector<T> rr(static_cast<size_t>(nstep) * (a.nspec - a.nspec_none) * NGLL * NGLL * 2);
      const int nsnap = snapshot_every > 0 ? (nstep - 2) / snapshot_every + 1 : 0;
      std::vector<T> ss(static_cast<size_t>(nsnap) * 6 * a.nglob);
      run_forward<T>(a, nsrc, S.data(), nrec, R.data(), nstep, dt, nsb, ntypes, types, seis, t0, nthreads,
                     ff.data(), rr.data(), snapshot_every, ss.data());
      for (size_t i = 0; i < ff.size(); i++) final_fields[i] = ff[i];
      for (size_t i = 0; i < rr.size(); i++) record[i] = rr[i];
      for (size_t i = 0; i < ss.size(); i++) snapshots[i] = ss[i];
    });
  });
}

// Combined adjoint + backward run.  adj_pos (nadj, 3) [x, z, angle];
// adj_stf (nadj, nstep, 2); kernels out (6, nspec, 5, 5) in mesh element
// order: rho, mu, kappa, rhop, alpha, beta; seis (nsig, ntypes, nrec, 2) of
// the backward field, latest sample first.
int orc2d_combined(void *h, int nsrc, const double *src, int nadj, const double *adj_pos, const double *adj_stf,
                   int nrec, const double *rec, int nstep, double dt, int nsb, int ntypes, const int *types,
                   const double *final_fields, const double *record, double *kernels, double *seis, int nthreads,
                   int snapshot_every, const double *snapshots) {
  return guard([&] {
    auto S = sources_in(nsrc, src);
    auto A = receivers_in(nadj, adj_pos);
    auto R = receivers_in(nrec, rec);
    DISPATCH(h, {
      auto &a = asm_of<T>(h);
      if (a.coupled()) throw std::invalid_argument("elastic meshes only");
      std::vector<T> ff(final_fields, final_fields + 6 * static_cast<size_t>(a.nglob));
      std::vector<T> rr(record, record + static_cast<size_t>(nstep) * (a.nspec - a.nspec_none) * NGLL * NGLL * 2);
      const size_t npts = static_cast<size_t>(a.nspec) * NGLL * NGLL;
      std::vector<T> K(6 * npts, 0);
      const int nsnap = snapshot_every > 0 ? (nstep - 2) / snapshot_every + 1 : 0;
      std::vector<T> ss(snapshots, snapshots + static_cast<size_t>(nsnap) * 6 * a.nglob);
      run_combined<T>(a, nsrc, S.data(), nadj, A.data(), adj_stf, nrec, R.data(), nstep, dt, nsb, ntypes, types,
                      ff.data(), rr.data(), K.data(), seis, nthreads, snapshot_every, ss.data());
      for (int k = 0; k < 6; k++)
        for (int c = 0; c < a.nspec; c++)
          for (int p = 0; p < NGLL * NGLL; p++)
            kernels[(k * static_cast<size_t>(a.nspec) + a.compute_to_mesh[c]) * NGLL * NGLL + p] =
                K[k * npts + static_cast<size_t>(c) * NGLL * NGLL + p];
    });
  });
}

// out (2, nglob) = assembled -K u (+ Stacey driven by v); u, v (2, nglob)
void orc2d_apply_stiffness(void *h, const double *u, const double *v, double *out) {
  DISPATCH(h, {
    auto &a = asm_of<T>(h);
    const size_t n = 2 * static_cast<size_t>(a.nglob);
    std::vector<T> uu(u, u + n), vv(v, v + n), oo(n);
    apply_stiffness<T>(a, uu.data(), vv.data(), oo.data());
    for (size_t i = 0; i < n; i++) out[i] = oo[i];
  });
}

}  // extern "C"
