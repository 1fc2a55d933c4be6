// ORACLE — test infrastructure only.
//
// CPU restatement of the 3-D elastic isotropic Newmark path on a structured
// hexahedral mesh.  The reference has no 3-D solver (its time_marching is
// instantiated for dim2 only, src/solver/time_marching.cpp:8-18), so this
// restatement extends the 2-D reference formulas to (xi, eta, gamma):
//   gradient ........ include/algorithms/gradient.hpp:81-103, one more direction
//   stress .......... isotropic Hooke law with lambda + 2 mu = kappa + 4/3 mu
//                     (core/specfem/point/properties.hpp:141-161)
//   F = J sigma d ... core/specfem/point/stress.hpp:118-139, 3x3
//   divergence ...... include/algorithms/divergence.hpp:86-98, weights of the
//                     two other directions
//   Stacey .......... include/boundary_conditions/stacey/stacey.tpp:36-72 with
//                     3-component vectors, per absorbing face
//   mass ............ compute_mass_matrix.tpp:80-114 + stacey.tpp:641-677
//   Newmark ......... include/timescheme/newmark.tpp:11-116
// Geometry is evaluated in general form from the 8-node trilinear map of each
// element (9 derivatives and the jacobian per point), independently of the
// engine's regular-element shortcut.
//
// Pinning: dense brute-force weak-form assembly on tiny meshes (symmetry,
// rigid-body null space, K u match; tests/test_oracle_bruteforce.py); in
// tests/test_oracle_3d.py the discrete central-difference energy, conserved
// to round-off without absorbing faces and balanced by the closed-form Stacey
// damping with them, the closed-form mass and Stacey mass term, the
// force-dipole limit of the moment-tensor source, and the threaded loop
// giving the serial loop's bits.  The numbering and element location follow
// the engine's documented model, not a reference artifact: "parity unpinned"
// against the reference for those.
#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc3 {

// Threads of the forward loop (1: the serial element loop).
int g_threads3 = 1;

constexpr int N = 5;

template <class T> T leg(T z, int n) {
  T p1 = 1.0, p2 = z, p3 = p2;
  for (int k = 1; k < n; k++) {
    double dk = k;
    p3 = ((2.0 * dk + 1.0) * z * p2 - dk * p1) / (dk + 1.0);
    p1 = p2;
    p2 = p3;
  }
  return p3;
}

template <class T> struct Quad {
  T xi[N], w[N], hp[N][N];  // hp[i][j] = l_j'(xi_i)
};

template <class T> Quad<T> quad() {
  // closed form for N = 5: 0, +-sqrt(3/7), +-1; weights 1/10, 49/90, 32/45
  Quad<T> q;
  const double x[N] = {-1.0, -std::sqrt(3.0 / 7.0), 0.0, std::sqrt(3.0 / 7.0), 1.0};
  const double w[N] = {0.1, 49.0 / 90.0, 32.0 / 45.0, 49.0 / 90.0, 0.1};
  for (int i = 0; i < N; i++) {
    q.xi[i] = static_cast<T>(x[i]);
    q.w[i] = static_cast<T>(w[i]);
  }
  // l_j'(x_i) = P4(x_i) / (P4(x_j) (x_i - x_j)), i != j; -/+ 5 at the ends
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      double v;
      if (i == 0 && j == 0) v = -5.0;
      else if (i == N - 1 && j == N - 1) v = 5.0;
      else if (i == j) v = 0.0;
      else v = leg<double>(x[i], 4) / (leg<double>(x[j], 4) * (x[i] - x[j]));
      q.hp[i][j] = static_cast<T>(v);
    }
  return q;
}

template <class T> void lagrange(T xi, const T *nodes, T *h, T *hp) {
  for (int d = 0; d < N; d++) {
    T p1 = 1.0, p2 = 1.0, sum = 0.0;
    for (int i = 0; i < N; i++) {
      if (i == d) continue;
      p1 *= (xi - nodes[i]);
      p2 *= (nodes[d] - nodes[i]);
      T p3 = 1.0;
      for (int j = 0; j < N; j++)
        if (j != d && j != i) p3 *= (xi - nodes[j]);
      sum += p3;
    }
    h[d] = p1 / p2;
    if (hp) hp[d] = sum / p2;
  }
}

template <class T> struct Asm {
  int nx, ny, nz, nspec, nglob;
  double x0, y0, z0, hx, hy, hz;
  Quad<T> q;
  std::vector<int> ibool;          // (nspec, 125), p = (k*5 + j)*5 + i
  std::vector<T> dxi;              // (nspec*125, 9): dxi/dx, dxi/dy, dxi/dz, deta/dx, ..., dgamma/dz
  std::vector<T> jac;              // (nspec*125)
  std::vector<T> kappa, mu, rho;   // (nspec*125)
  struct Bnd {
    int e, p;
    T n[3], w, rvp, rvs;
  };
  std::vector<Bnd> bnd;
  std::vector<T> minv;             // (3, nglob)
  // mesh databases (general hexahedra): the GLL point coordinates, read by
  // the general locator
  bool general = false;
  std::vector<double> px, py, pz;  // (nglob)
};

// Trilinear map of a box element evaluated in general form.
template <class T>
void element_geometry(const Asm<T> &a, int ex, int ey, int ez, int i, int j, int k, T *d9, T &J) {
  // corner coordinates
  T cx[8], cy[8], cz[8];
  for (int c = 0; c < 8; c++) {
    const int ix = c & 1, iy = (c >> 1) & 1, iz = (c >> 2) & 1;
    cx[c] = static_cast<T>(a.x0 + (ex + ix) * a.hx);
    cy[c] = static_cast<T>(a.y0 + (ey + iy) * a.hy);
    cz[c] = static_cast<T>(a.z0 + (ez + iz) * a.hz);
  }
  const T xi = a.q.xi[i], et = a.q.xi[j], ga = a.q.xi[k];
  T dX[3][3] = {{0}};  // dX[r][s] = d x_r / d xi_s
  for (int c = 0; c < 8; c++) {
    const T sx = (c & 1) ? 1 : -1, sy = ((c >> 1) & 1) ? 1 : -1, sz = ((c >> 2) & 1) ? 1 : -1;
    const T fx = (1 + sx * xi), fy = (1 + sy * et), fz = (1 + sz * ga);
    const T dN[3] = {sx * fy * fz / 8, fx * sy * fz / 8, fx * fy * sz / 8};
    const T X[3] = {cx[c], cy[c], cz[c]};
    for (int r = 0; r < 3; r++)
      for (int s = 0; s < 3; s++) dX[r][s] += X[r] * dN[s];
  }
  J = dX[0][0] * (dX[1][1] * dX[2][2] - dX[1][2] * dX[2][1]) - dX[0][1] * (dX[1][0] * dX[2][2] - dX[1][2] * dX[2][0]) +
      dX[0][2] * (dX[1][0] * dX[2][1] - dX[1][1] * dX[2][0]);
  // inverse: d xi_s / d x_r = cof / J, stored as d9[s*3 + r]
  T inv[3][3];
  inv[0][0] = (dX[1][1] * dX[2][2] - dX[1][2] * dX[2][1]) / J;
  inv[0][1] = (dX[0][2] * dX[2][1] - dX[0][1] * dX[2][2]) / J;
  inv[0][2] = (dX[0][1] * dX[1][2] - dX[0][2] * dX[1][1]) / J;
  inv[1][0] = (dX[1][2] * dX[2][0] - dX[1][0] * dX[2][2]) / J;
  inv[1][1] = (dX[0][0] * dX[2][2] - dX[0][2] * dX[2][0]) / J;
  inv[1][2] = (dX[0][2] * dX[1][0] - dX[0][0] * dX[1][2]) / J;
  inv[2][0] = (dX[1][0] * dX[2][1] - dX[1][1] * dX[2][0]) / J;
  inv[2][1] = (dX[0][1] * dX[2][0] - dX[0][0] * dX[2][1]) / J;
  inv[2][2] = (dX[0][0] * dX[1][1] - dX[0][1] * dX[1][0]) / J;
  for (int s = 0; s < 3; s++)
    for (int r = 0; r < 3; r++) d9[s * 3 + r] = inv[s][r];
}

template <class T>
void stacey3(const T *n, T w, T rvp, T rvs, const T *v, T *out) {
  T vn = 0, nn = 0;
  for (int d = 0; d < 3; d++) {
    vn += v[d] * n[d];
    nn += n[d] * n[d];
  }
  const T j1 = std::sqrt(nn);
  for (int d = 0; d < 3; d++) {
    const T f = ((vn * n[d] / (j1 * j1)) * (rvp - rvs)) + v[d] * rvs;
    out[d] += static_cast<T>(-1.0) * f * j1 * w;
  }
}

template <class T>
Asm<T> *build(int nx, int ny, int nz, double x0, double y0, double z0, double hx, double hy, double hz,
              const int *material, const double *rho, const double *vp, const double *vs, const int *absorbing,
              double dt) {
  auto *a = new Asm<T>();
  a->nx = nx; a->ny = ny; a->nz = nz;
  a->x0 = x0; a->y0 = y0; a->z0 = z0; a->hx = hx; a->hy = hy; a->hz = hz;
  a->nspec = nx * ny * nz;
  a->q = quad<T>();
  const int nspec = a->nspec;
  // numbering: first visit in (element, k, j, i) order; shared points are
  // identified by exact lattice position
  const long LX = 4L * nx + 1, LY = 4L * ny + 1, LZ = 4L * nz + 1;
  std::vector<int> first(static_cast<size_t>(LX * LY * LZ), -1);
  a->ibool.resize(static_cast<size_t>(nspec) * 125);
  int inum = 0;
  for (int e = 0; e < nspec; e++) {
    const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const long L = ((4L * ez + k) * LY + (4L * ey + j)) * LX + (4L * ex + i);
          int &f = first[L];
          if (f < 0) f = inum++;
          a->ibool[static_cast<size_t>(e) * 125 + (k * 5 + j) * 5 + i] = f;
        }
  }
  a->nglob = inum;
  const size_t np = static_cast<size_t>(nspec) * 125;
  a->dxi.resize(np * 9);
  a->jac.resize(np);
  a->kappa.resize(np);
  a->mu.resize(np);
  a->rho.resize(np);
  for (int e = 0; e < nspec; e++) {
    const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    const int m = material[e];
    const T r = static_cast<T>(rho[m]), cp = static_cast<T>(vp[m]), cs = static_cast<T>(vs[m]);
    const T mu = r * cs * cs;
    const T kappa = r * (cp * cp - (4.0 / 3.0) * cs * cs);
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const size_t p = static_cast<size_t>(e) * 125 + (k * 5 + j) * 5 + i;
          element_geometry(*a, ex, ey, ez, i, j, k, &a->dxi[p * 9], a->jac[p]);
          a->kappa[p] = kappa;
          a->mu[p] = mu;
          a->rho[p] = r;
        }
  }
  // Stacey faces: outward normal n = J * (grad of the face-normal reference
  // coordinate) with the sign of the face, weight = in-face GLL weights.
  for (int e = 0; e < nspec; e++) {
    const int ex = e % nx, ey = (e / nx) % ny, ez = e / (nx * ny);
    for (int face = 0; face < 6; face++) {
      if (!absorbing[face]) continue;
      const bool on = (face == 0 && ex == 0) || (face == 1 && ex == nx - 1) || (face == 2 && ey == 0) ||
                      (face == 3 && ey == ny - 1) || (face == 4 && ez == 0) || (face == 5 && ez == nz - 1);
      if (!on) continue;
      const int dir = face / 2;
      const T sgn = (face % 2) ? 1 : -1;
      for (int b = 0; b < N; b++)
        for (int c = 0; c < N; c++) {
          int idx[3];
          idx[dir] = (face % 2) ? N - 1 : 0;
          const int o1 = (dir + 1) % 3, o2 = (dir + 2) % 3;
          // in-face indices in ascending axis order
          const int lo = std::min(o1, o2), hi = std::max(o1, o2);
          idx[lo] = c;
          idx[hi] = b;
          const int i = idx[0], j = idx[1], k = idx[2];
          const int p = (k * 5 + j) * 5 + i;
          const size_t pp = static_cast<size_t>(e) * 125 + p;
          typename Asm<T>::Bnd B;
          B.e = e;
          B.p = p;
          for (int d = 0; d < 3; d++) B.n[d] = sgn * a->jac[pp] * a->dxi[pp * 9 + dir * 3 + d];
          B.w = a->q.w[idx[lo]] * a->q.w[idx[hi]];
          const T lp2m = a->kappa[pp] + static_cast<T>(4.0 / 3.0) * a->mu[pp];
          B.rvp = std::sqrt(lp2m / a->rho[pp]) * a->rho[pp];
          B.rvs = std::sqrt(a->rho[pp] * a->mu[pp]);
          a->bnd.push_back(B);
        }
    }
  }
  // mass
  std::vector<T> mass(3 * static_cast<size_t>(a->nglob), 0);
  for (int e = 0; e < nspec; e++)
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const size_t p = static_cast<size_t>(e) * 125 + (k * 5 + j) * 5 + i;
          const T mm = a->rho[p] * (a->q.w[i] * a->q.w[j] * a->q.w[k] * a->jac[p]);
          const int g = a->ibool[p];
          for (int c = 0; c < 3; c++) mass[static_cast<size_t>(c) * a->nglob + g] += mm;
        }
  const T vf = static_cast<T>(-1.0) * static_cast<T>(dt) * static_cast<T>(0.5);
  for (const auto &B : a->bnd) {
    T t[3] = {0, 0, 0};
    const T v[3] = {vf, vf, vf};
    stacey3(B.n, B.w, B.rvp, B.rvs, v, t);
    const int g = a->ibool[static_cast<size_t>(B.e) * 125 + B.p];
    for (int c = 0; c < 3; c++) mass[static_cast<size_t>(c) * a->nglob + g] += t[c];
  }
  a->minv.resize(mass.size());
  for (size_t i = 0; i < mass.size(); i++) a->minv[i] = static_cast<T>(1.0) / mass[i];
  return a;
}

// A 3-D mesh database (src/io/mesh/impl/fortran/dim3/mesh.cpp:43-797; no 3-D
// solver exists in the reference): numbering as given, per-point derivatives
// of the irregular elements, xix_regular / jacobian_regular for the others
// (mesh.cpp:150-151), per-point properties, absorbing faces with unit normal
// and jacobian2Dw (the traction of stacey3 with |n| = 1), mass as in build().
template <class T>
Asm<T> *build_db(int nspec, int nglob, const int *ibool, const float *x, const float *y, const float *z, int nirr,
                 const int *irr, float xix_reg, float jac_reg, const float *const *d9in, const float *jac,
                 const float *kappa, const float *mu, const float *rho, const float *rvp, const float *rvs, int nabs,
                 const int *abs_ispec, const int *abs_ijk, const float *abs_jw, const float *abs_n, double dt) {
  auto *a = new Asm<T>();
  a->general = true;
  a->nx = a->ny = a->nz = 0;
  a->nspec = nspec;
  a->nglob = nglob;
  a->q = quad<T>();
  const size_t np = static_cast<size_t>(nspec) * 125;
  a->ibool.assign(ibool, ibool + np);
  for (int v : a->ibool)
    if (v < 0 || v >= nglob) throw std::runtime_error("ibool out of range");
  a->px.assign(x, x + nglob);
  a->py.assign(y, y + nglob);
  a->pz.assign(z, z + nglob);
  std::vector<char> irregular(nspec, 0);
  for (int i = 0; i < nirr; i++) irregular[irr[i]] = 1;
  a->dxi.assign(np * 9, 0);
  a->jac.resize(np);
  a->kappa.resize(np);
  a->mu.resize(np);
  a->rho.resize(np);
  for (size_t p = 0; p < np; p++) {
    const int e = static_cast<int>(p / 125);
    if (irregular[e]) {
      for (int d = 0; d < 9; d++) a->dxi[p * 9 + d] = static_cast<T>(d9in[d][p]);
      a->jac[p] = static_cast<T>(jac[p]);
    } else {
      a->dxi[p * 9 + 0] = a->dxi[p * 9 + 4] = a->dxi[p * 9 + 8] = static_cast<T>(xix_reg);
      a->jac[p] = static_cast<T>(jac_reg);
    }
    a->kappa[p] = static_cast<T>(kappa[p]);
    a->mu[p] = static_cast<T>(mu[p]);
    a->rho[p] = static_cast<T>(rho[p]);
  }
  for (int f = 0; f < nabs; f++)
    for (int qf = 0; qf < 25; qf++) {
      const int i = abs_ijk[(f * 3 + 0) * 25 + qf], j = abs_ijk[(f * 3 + 1) * 25 + qf],
                k = abs_ijk[(f * 3 + 2) * 25 + qf];
      typename Asm<T>::Bnd B;
      B.e = abs_ispec[f];
      B.p = (k * 5 + j) * 5 + i;
      for (int d = 0; d < 3; d++) B.n[d] = static_cast<T>(abs_n[(f * 3 + d) * 25 + qf]);
      B.w = static_cast<T>(abs_jw[f * 25 + qf]);
      const size_t pp = static_cast<size_t>(B.e) * 125 + B.p;
      B.rvp = static_cast<T>(rvp[pp]);
      B.rvs = static_cast<T>(rvs[pp]);
      a->bnd.push_back(B);
    }
  std::vector<T> mass(3 * static_cast<size_t>(nglob), 0);
  for (int e = 0; e < nspec; e++)
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const size_t p = static_cast<size_t>(e) * 125 + (k * 5 + j) * 5 + i;
          const T mm = a->rho[p] * (a->q.w[i] * a->q.w[j] * a->q.w[k] * a->jac[p]);
          for (int c = 0; c < 3; c++) mass[static_cast<size_t>(c) * nglob + a->ibool[p]] += mm;
        }
  const T vf = static_cast<T>(-1.0) * static_cast<T>(dt) * static_cast<T>(0.5);
  for (const auto &B : a->bnd) {
    T t[3] = {0, 0, 0};
    const T v[3] = {vf, vf, vf};
    stacey3(B.n, B.w, B.rvp, B.rvs, v, t);
    const int g = a->ibool[static_cast<size_t>(B.e) * 125 + B.p];
    for (int c = 0; c < 3; c++) mass[static_cast<size_t>(c) * nglob + g] += t[c];
  }
  a->minv.resize(mass.size());
  for (size_t i = 0; i < mass.size(); i++) a->minv[i] = static_cast<T>(1.0) / mass[i];
  return a;
}

// Point location on a mesh database (the engine's model; the reference has
// no 3-D locator): the nearest GLL point (double squared distance, first
// minimum in global order), then in each element holding it (ascending) 100
// Newton steps on the isoparametric map of the element's GLL points, clamped
// to [-1.01, 1.01], from that point's reference coordinates; the smallest
// final distance wins, first on ties.  Computed in double.
template <class T> void locate_general(const Asm<T> &a, double x, double y, double z, int &ispec, double r_out[3]) {
  const double xn[N] = {-1.0, -std::sqrt(3.0 / 7.0), 0.0, std::sqrt(3.0 / 7.0), 1.0};
  int best = -1;
  double dmin = std::numeric_limits<double>::max();
  for (int g = 0; g < a.nglob; g++) {
    const double d = (a.px[g] - x) * (a.px[g] - x) + (a.py[g] - y) * (a.py[g] - y) + (a.pz[g] - z) * (a.pz[g] - z);
    if (d < dmin) {
      dmin = d;
      best = g;
    }
  }
  auto map = [&](int e, const double r[3], double X[3], double D[3][3]) {
    double h[3][N], hp[3][N];
    for (int d = 0; d < 3; d++) lagrange<double>(r[d], xn, h[d], hp[d]);
    for (int c = 0; c < 3; c++) {
      X[c] = 0;
      for (int s2 = 0; s2 < 3; s2++) D[c][s2] = 0;
    }
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const int g = a.ibool[static_cast<size_t>(e) * 125 + (k * 5 + j) * 5 + i];
          const double P[3] = {a.px[g], a.py[g], a.pz[g]};
          const double w = h[0][i] * h[1][j] * h[2][k];
          const double dw[3] = {hp[0][i] * h[1][j] * h[2][k], h[0][i] * hp[1][j] * h[2][k], h[0][i] * h[1][j] * hp[2][k]};
          for (int c = 0; c < 3; c++) {
            X[c] += w * P[c];
            for (int s2 = 0; s2 < 3; s2++) D[c][s2] += dw[s2] * P[c];
          }
        }
  };
  double final_best = std::numeric_limits<double>::max();
  ispec = -1;
  struct Cand {
    int e;
    double d, r[3];
  };
  std::vector<Cand> cand;
  for (int e = 0; e < a.nspec; e++) {
    int pbest = -1;
    for (int p = 0; p < 125 && pbest < 0; p++)
      if (a.ibool[static_cast<size_t>(e) * 125 + p] == best) pbest = p;
    if (pbest < 0) continue;
    double r[3] = {xn[pbest % 5], xn[(pbest / 5) % 5], xn[pbest / 25]};
    double X[3], D[3][3];
    for (int it = 0; it < 100; it++) {
      map(e, r, X, D);
      const double res[3] = {x - X[0], y - X[1], z - X[2]};
      const double det = D[0][0] * (D[1][1] * D[2][2] - D[1][2] * D[2][1]) -
                         D[0][1] * (D[1][0] * D[2][2] - D[1][2] * D[2][0]) +
                         D[0][2] * (D[1][0] * D[2][1] - D[1][1] * D[2][0]);
      if (det == 0.0) break;
      double inv[3][3];
      inv[0][0] = (D[1][1] * D[2][2] - D[1][2] * D[2][1]) / det;
      inv[0][1] = (D[0][2] * D[2][1] - D[0][1] * D[2][2]) / det;
      inv[0][2] = (D[0][1] * D[1][2] - D[0][2] * D[1][1]) / det;
      inv[1][0] = (D[1][2] * D[2][0] - D[1][0] * D[2][2]) / det;
      inv[1][1] = (D[0][0] * D[2][2] - D[0][2] * D[2][0]) / det;
      inv[1][2] = (D[0][2] * D[1][0] - D[0][0] * D[1][2]) / det;
      inv[2][0] = (D[1][0] * D[2][1] - D[1][1] * D[2][0]) / det;
      inv[2][1] = (D[0][1] * D[2][0] - D[0][0] * D[2][1]) / det;
      inv[2][2] = (D[0][0] * D[1][1] - D[0][1] * D[1][0]) / det;
      for (int s2 = 0; s2 < 3; s2++)
        r[s2] = std::min(1.01, std::max(-1.01, r[s2] + inv[s2][0] * res[0] + inv[s2][1] * res[1] + inv[s2][2] * res[2]));
    }
    map(e, r, X, D);
    const double d = std::sqrt((x - X[0]) * (x - X[0]) + (y - X[1]) * (y - X[1]) + (z - X[2]) * (z - X[2]));
    cand.push_back({e, d, {r[0], r[1], r[2]}});
    final_best = std::min(final_best, d);
  }
  // ties (a point on a shared face, edge or corner: distances equal up to
  // rounding, within 1e-9 of the mesh's bounding-box diagonal) go to the
  // highest-numbered element, as in the structured location rule below
  double lo[3] = {a.px[0], a.py[0], a.pz[0]}, hi[3] = {a.px[0], a.py[0], a.pz[0]};
  for (int g = 0; g < a.nglob; g++) {
    lo[0] = std::min(lo[0], static_cast<double>(a.px[g]));
    lo[1] = std::min(lo[1], static_cast<double>(a.py[g]));
    lo[2] = std::min(lo[2], static_cast<double>(a.pz[g]));
    hi[0] = std::max(hi[0], static_cast<double>(a.px[g]));
    hi[1] = std::max(hi[1], static_cast<double>(a.py[g]));
    hi[2] = std::max(hi[2], static_cast<double>(a.pz[g]));
  }
  const double tie = 1e-9 * std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                      (hi[2] - lo[2]) * (hi[2] - lo[2]));
  for (const auto &c : cand)
    if (c.d <= final_best + tie && c.e > ispec) {
      ispec = c.e;
      for (int s2 = 0; s2 < 3; s2++) r_out[s2] = c.r[s2];
    }
  if (ispec < 0) throw std::runtime_error("locate: no element holds the nearest point");
}

// -K_e u for one element (plus nothing else), general geometry.
template <class T> void element_force(const Asm<T> &a, int e, const T *u /*(3,nglob)*/, T out[125][3]) {
  const auto &q = a.q;
  const int ng = a.nglob;
  T ul[125][3];
  for (int p = 0; p < 125; p++) {
    const int g = a.ibool[static_cast<size_t>(e) * 125 + p];
    for (int c = 0; c < 3; c++) ul[p][c] = u[static_cast<size_t>(c) * ng + g];
  }
  static thread_local T F[125][3][3];
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const int p = (k * 5 + j) * 5 + i;
        T dref[3][3] = {{0}};  // dref[c][s] = d u_c / d xi_s
        for (int l = 0; l < N; l++)
          for (int c = 0; c < 3; c++) {
            dref[c][0] += q.hp[i][l] * ul[(k * 5 + j) * 5 + l][c];
            dref[c][1] += q.hp[j][l] * ul[(k * 5 + l) * 5 + i][c];
            dref[c][2] += q.hp[k][l] * ul[(l * 5 + j) * 5 + i][c];
          }
        const size_t pp = static_cast<size_t>(e) * 125 + p;
        const T *d9 = &a.dxi[pp * 9];
        T du[3][3];  // du[c][r] = d u_c / d x_r
        for (int c = 0; c < 3; c++)
          for (int r = 0; r < 3; r++) du[c][r] = dref[c][0] * d9[0 * 3 + r] + dref[c][1] * d9[1 * 3 + r] + dref[c][2] * d9[2 * 3 + r];
        const T mu = a.mu[pp];
        const T lp2m = a.kappa[pp] + static_cast<T>(4.0 / 3.0) * mu;
        const T lam = lp2m - static_cast<T>(2.0) * mu;
        const T tr = du[0][0] + du[1][1] + du[2][2];
        T s[3][3];
        for (int c = 0; c < 3; c++)
          for (int r = 0; r < 3; r++) s[c][r] = mu * (du[c][r] + du[r][c]) + (c == r ? lam * tr : static_cast<T>(0));
        const T J = a.jac[pp];
        for (int c = 0; c < 3; c++)
          for (int sdir = 0; sdir < 3; sdir++)
            F[p][c][sdir] = J * (s[c][0] * d9[sdir * 3 + 0] + s[c][1] * d9[sdir * 3 + 1] + s[c][2] * d9[sdir * 3 + 2]);
      }
  for (int k = 0; k < N; k++)
    for (int j = 0; j < N; j++)
      for (int i = 0; i < N; i++) {
        const int p = (k * 5 + j) * 5 + i;
        for (int c = 0; c < 3; c++) {
          T t0 = 0, t1 = 0, t2 = 0;
          for (int l = 0; l < N; l++) {
            t0 += F[(k * 5 + j) * 5 + l][c][0] * q.hp[l][i] * q.w[l];
            t1 += F[(k * 5 + l) * 5 + i][c][1] * q.hp[l][j] * q.w[l];
            t2 += F[(l * 5 + j) * 5 + i][c][2] * q.hp[l][k] * q.w[l];
          }
          out[p][c] = -(q.w[j] * q.w[k] * t0 + q.w[i] * q.w[k] * t1 + q.w[i] * q.w[j] * t2);
        }
      }
}

template <class T>
void forward(const Asm<T> &a, int nsrc, const double *src /*(nsrc,16)*/, int nrec, const double *rec /*(nrec,3)*/,
             int nstep, double dt_in, int nsb, int ntypes, const int *types, double *seis, double *t0_out,
             double *final_fields) {
  const int ng = a.nglob;
  const T dt = static_cast<T>(dt_in), dto2 = dt / 2.0, dt2o2 = dt * dt / 2.0;
  // sources: rows [kind, x, y, z, fx, fy, fz, mxx, myy, mzz, mxy, mxz, myz, f0, tshift, factor]
  T min_t0 = std::numeric_limits<T>::max();
  std::vector<T> t0s(nsrc);
  for (int s = 0; s < nsrc; s++) {
    const T hdur = 1.0 / static_cast<T>(src[16 * s + 13]);
    t0s[s] = -1.2 * hdur + static_cast<T>(src[16 * s + 14]);
    min_t0 = std::min(min_t0, t0s[s]);
  }
  if (nsrc == 0) min_t0 = 0;
  *t0_out = min_t0;
  auto locate = [&](double x, double y, double z, int &e, T &xi, T &et, T &ga) {
    if (a.general) {
      double r[3];
      locate_general<T>(a, x, y, z, e, r);
      xi = static_cast<T>(r[0]);
      et = static_cast<T>(r[1]);
      ga = static_cast<T>(r[2]);
      return;
    }
    auto one = [](double v, double v0, double h, int n, int &k, double &r) {
      const double s = (v - v0) / h;
      k = std::min(std::max(static_cast<int>(std::floor(s)), 0), n - 1);
      r = std::min(std::max(2.0 * (s - k) - 1.0, -1.0), 1.0);
    };
    int ex, ey, ez;
    double rx, ry, rz;
    one(x, a.x0, a.hx, a.nx, ex, rx);
    one(y, a.y0, a.hy, a.ny, ey, ry);
    one(z, a.z0, a.hz, a.nz, ez, rz);
    e = (ez * a.ny + ey) * a.nx + ex;
    xi = static_cast<T>(rx);
    et = static_cast<T>(ry);
    ga = static_cast<T>(rz);
  };
  struct Src {
    int e;
    T arr[3][125];
    std::vector<T> stf;
  };
  std::vector<Src> S(nsrc);
  for (int s = 0; s < nsrc; s++) {
    const double *r = &src[16 * s];
    T xi, et, ga;
    locate(r[1], r[2], r[3], S[s].e, xi, et, ga);
    T hx[N], hpx[N], hy[N], hpy[N], hz[N], hpz[N];
    lagrange<T>(xi, a.q.xi, hx, hpx);
    lagrange<T>(et, a.q.xi, hy, hpy);
    lagrange<T>(ga, a.q.xi, hz, hpz);
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) {
          const int p = (k * 5 + j) * 5 + i;
          if (static_cast<int>(r[0]) == 0) {
            const T h = hx[i] * hy[j] * hz[k];
            for (int c = 0; c < 3; c++) S[s].arr[c][p] = static_cast<T>(r[4 + c]) * h;
          } else {
            // gradient of the Lagrange product through the element geometry:
            // on boxes the (constant) centre value; on a mesh database the
            // derivatives averaged with the source point's Lagrange weights
            // (the 3-D form of moment_tensor_source.cpp:88-130)
            const size_t pp = static_cast<size_t>(S[s].e) * 125 + 62;  // element centre (constant on boxes)
            T dav[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (a.general)
              for (int kk = 0; kk < N; kk++)
                for (int jj = 0; jj < N; jj++)
                  for (int ii = 0; ii < N; ii++) {
                    const T hl = hx[ii] * hy[jj] * hz[kk];
                    const size_t q2 = static_cast<size_t>(S[s].e) * 125 + (kk * 5 + jj) * 5 + ii;
                    for (int d = 0; d < 9; d++) dav[d] += hl * a.dxi[q2 * 9 + d];
                  }
            const T *d9 = a.general ? dav : &a.dxi[pp * 9];
            const T dref[3] = {hpx[i] * hy[j] * hz[k], hx[i] * hpy[j] * hz[k], hx[i] * hy[j] * hpz[k]};
            T dx[3];
            for (int rr = 0; rr < 3; rr++) dx[rr] = dref[0] * d9[0 * 3 + rr] + dref[1] * d9[1 * 3 + rr] + dref[2] * d9[2 * 3 + rr];
            const T M[3][3] = {{static_cast<T>(r[7]), static_cast<T>(r[10]), static_cast<T>(r[11])},
                               {static_cast<T>(r[10]), static_cast<T>(r[8]), static_cast<T>(r[12])},
                               {static_cast<T>(r[11]), static_cast<T>(r[12]), static_cast<T>(r[9])}};
            for (int c = 0; c < 3; c++) S[s].arr[c][p] = M[c][0] * dx[0] + M[c][1] * dx[1] + M[c][2] * dx[2];
          }
        }
    S[s].stf.resize(nstep);
    const T shift = t0s[s] - min_t0;
    const T f0 = static_cast<T>(r[13]), factor = static_cast<T>(r[15]);
    const T pi = static_cast<T>(M_PI);
    for (int it = 0; it < nstep; it++) {
      const T t = min_t0 + it * dt - shift;
      const T aa = pi * pi * f0 * f0;
      S[s].stf[it] = -1.0 * factor * ((1.0 - 2.0 * aa * t * t) * std::exp(-aa * t * t));
    }
  }
  struct Rec {
    int e;
    T h[125];
  };
  std::vector<Rec> R(nrec);
  for (int r = 0; r < nrec; r++) {
    T xi, et, ga;
    locate(rec[3 * r], rec[3 * r + 1], rec[3 * r + 2], R[r].e, xi, et, ga);
    T hx[N], hy[N], hz[N];
    lagrange<T>(xi, a.q.xi, hx, nullptr);
    lagrange<T>(et, a.q.xi, hy, nullptr);
    lagrange<T>(ga, a.q.xi, hz, nullptr);
    for (int k = 0; k < N; k++)
      for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) R[r].h[(k * 5 + j) * 5 + i] = hx[i] * hy[j] * hz[k];
  }
  std::vector<T> u(3 * static_cast<size_t>(ng), 0), v(u.size(), 0), acc(u.size(), 0);
  T out[125][3];
  // Threaded variant (timing baseline only): element forces of all elements
  // into ef in parallel, then each point adds its contributions in
  // increasing element order -- the order of the serial element loop, so
  // both give the same bits.
  const int nthreads = std::max(1, g_threads3);
  const long n3 = static_cast<long>(u.size());
  std::vector<T> ef;
  std::vector<int> inv_off, inv_ent;
  if (nthreads > 1) {
    ef.resize(static_cast<size_t>(a.nspec) * 375);
    inv_off.assign(static_cast<size_t>(ng) + 1, 0);
    for (size_t q = 0; q < a.ibool.size(); q++) inv_off[a.ibool[q] + 1]++;
    for (int g = 0; g < ng; g++) inv_off[g + 1] += inv_off[g];
    inv_ent.resize(a.ibool.size());
    std::vector<int> fill(inv_off.begin(), inv_off.end() - 1);
    for (size_t q = 0; q < a.ibool.size(); q++) inv_ent[fill[a.ibool[q]]++] = static_cast<int>(q);
  }
  int isig = 0;
  for (int istep = 0; istep < nstep; istep++) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long i = 0; i < n3; i++) {
      u[i] += dt * v[i] + dt2o2 * acc[i];
      v[i] += dto2 * acc[i];
      acc[i] = 0;
    }
    for (int s = 0; s < nsrc; s++)
      for (int p = 0; p < 125; p++) {
        const int g = a.ibool[static_cast<size_t>(S[s].e) * 125 + p];
        for (int c = 0; c < 3; c++) acc[static_cast<size_t>(c) * ng + g] += S[s].stf[istep] * S[s].arr[c][p];
      }
    if (nthreads > 1) {
#pragma omp parallel for schedule(static) num_threads(nthreads)
      for (int e = 0; e < a.nspec; e++)
        element_force<T>(a, e, u.data(), reinterpret_cast<T(*)[3]>(&ef[static_cast<size_t>(e) * 375]));
#pragma omp parallel for schedule(static) num_threads(nthreads)
      for (int g = 0; g < ng; g++)
        for (int q = inv_off[g]; q < inv_off[g + 1]; q++)
          for (int c = 0; c < 3; c++) acc[static_cast<size_t>(c) * ng + g] += ef[static_cast<size_t>(inv_ent[q]) * 3 + c];
    } else {
      for (int e = 0; e < a.nspec; e++) {
        element_force<T>(a, e, u.data(), out);
        for (int p = 0; p < 125; p++) {
          const int g = a.ibool[static_cast<size_t>(e) * 125 + p];
          for (int c = 0; c < 3; c++) acc[static_cast<size_t>(c) * ng + g] += out[p][c];
        }
      }
    }
    for (const auto &B : a.bnd) {
      const int g = a.ibool[static_cast<size_t>(B.e) * 125 + B.p];
      T vv[3], t[3] = {0, 0, 0};
      for (int c = 0; c < 3; c++) vv[c] = v[static_cast<size_t>(c) * ng + g];
      stacey3(B.n, B.w, B.rvp, B.rvs, vv, t);
      for (int c = 0; c < 3; c++) acc[static_cast<size_t>(c) * ng + g] += t[c];
    }
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long i = 0; i < n3; i++) {
      acc[i] = acc[i] * a.minv[i];
      v[i] += dto2 * acc[i];
    }
    if (istep % nsb == 0) {
      for (int it = 0; it < ntypes; it++) {
        const std::vector<T> &fld = types[it] == 0 ? u : (types[it] == 1 ? v : acc);
        for (int r = 0; r < nrec; r++) {
          T sc[3] = {0, 0, 0};
          for (int p = 0; p < 125; p++) {
            const int g = a.ibool[static_cast<size_t>(R[r].e) * 125 + p];
            for (int c = 0; c < 3; c++) sc[c] += R[r].h[p] * fld[static_cast<size_t>(c) * ng + g];
          }
          for (int c = 0; c < 3; c++) seis[((static_cast<size_t>(isig) * ntypes + it) * nrec + r) * 3 + c] = sc[c];
        }
      }
      isig++;
    }
  }
  if (final_fields) {
    for (size_t i = 0; i < u.size(); i++) {
      final_fields[i] = u[i];
      final_fields[u.size() + i] = v[i];
      final_fields[2 * u.size() + i] = acc[i];
    }
  }
}

template <class T> void apply_stiffness(const Asm<T> &a, const double *u, double *out) {
  const size_t n = 3 * static_cast<size_t>(a.nglob);
  std::vector<T> uu(u, u + n), oo(n, 0);
  T o[125][3];
  for (int e = 0; e < a.nspec; e++) {
    element_force<T>(a, e, uu.data(), o);
    for (int p = 0; p < 125; p++) {
      const int g = a.ibool[static_cast<size_t>(e) * 125 + p];
      for (int c = 0; c < 3; c++) oo[static_cast<size_t>(c) * a.nglob + g] += o[p][c];
    }
  }
  for (size_t i = 0; i < n; i++) out[i] = oo[i];
}

}  // namespace orc3

namespace {
thread_local std::string g_err3;
struct H3 {
  int precision;
  void *a;
};
}  // namespace

extern "C" {

const char *orc3d_last_error() { return g_err3.c_str(); }

// Threads used by orc3d_forward from now on (1: the serial loop; both give the
// same bits).
void orc3d_set_threads(int n) { orc3::g_threads3 = n < 1 ? 1 : n; }

int orc3d_create(int precision, int nx, int ny, int nz, double x0, double y0, double z0, double hx, double hy,
                 double hz, const int *material, const double *rho, const double *vp, const double *vs,
                 const int *absorbing, double dt, void **out) {
  try {
    auto *h = new H3{precision, nullptr};
    if (precision == 4)
      h->a = orc3::build<float>(nx, ny, nz, x0, y0, z0, hx, hy, hz, material, rho, vp, vs, absorbing, dt);
    else
      h->a = orc3::build<double>(nx, ny, nz, x0, y0, z0, hx, hy, hz, material, rho, vp, vs, absorbing, dt);
    *out = h;
    return 0;
  } catch (const std::exception &e) {
    g_err3 = e.what();
    return -1;
  }
}

int orc3d_create_db(int precision, int nspec, int nglob, const int *ibool, const float *x, const float *y,
                    const float *z, int nirr, const int *irr, float xix_reg, float jac_reg, const float *const *d9,
                    const float *jac, const float *kappa, const float *mu, const float *rho, const float *rvp,
                    const float *rvs, int nabs, const int *abs_ispec, const int *abs_ijk, const float *abs_jw,
                    const float *abs_n, double dt, void **out) {
  try {
    auto *h = new H3{precision, nullptr};
    if (precision == 4)
      h->a = orc3::build_db<float>(nspec, nglob, ibool, x, y, z, nirr, irr, xix_reg, jac_reg, d9, jac, kappa, mu, rho,
                                   rvp, rvs, nabs, abs_ispec, abs_ijk, abs_jw, abs_n, dt);
    else
      h->a = orc3::build_db<double>(nspec, nglob, ibool, x, y, z, nirr, irr, xix_reg, jac_reg, d9, jac, kappa, mu,
                                    rho, rvp, rvs, nabs, abs_ispec, abs_ijk, abs_jw, abs_n, dt);
    *out = h;
    return 0;
  } catch (const std::exception &e) {
    g_err3 = e.what();
    return -1;
  }
}

void orc3d_destroy(void *p) {
  auto *h = static_cast<H3 *>(p);
  if (!h) return;
  if (h->precision == 4)
    delete static_cast<orc3::Asm<float> *>(h->a);
  else
    delete static_cast<orc3::Asm<double> *>(h->a);
  delete h;
}

int orc3d_nglob(void *p) {
  auto *h = static_cast<H3 *>(p);
  return h->precision == 4 ? static_cast<orc3::Asm<float> *>(h->a)->nglob : static_cast<orc3::Asm<double> *>(h->a)->nglob;
}

void orc3d_ibool(void *p, int *out) {
  auto *h = static_cast<H3 *>(p);
  const auto &ib = h->precision == 4 ? static_cast<orc3::Asm<float> *>(h->a)->ibool
                                     : static_cast<orc3::Asm<double> *>(h->a)->ibool;
  std::copy(ib.begin(), ib.end(), out);
}

void orc3d_mass_inverse(void *p, double *out) {
  auto *h = static_cast<H3 *>(p);
  if (h->precision == 4) {
    const auto &m = static_cast<orc3::Asm<float> *>(h->a)->minv;
    for (size_t i = 0; i < m.size(); i++) out[i] = m[i];
  } else {
    const auto &m = static_cast<orc3::Asm<double> *>(h->a)->minv;
    for (size_t i = 0; i < m.size(); i++) out[i] = m[i];
  }
}

int orc3d_forward(void *p, int nsrc, const double *src, int nrec, const double *rec, int nstep, double dt, int nsb,
                  int ntypes, const int *types, double *seis, double *t0, double *final_fields) {
  try {
    auto *h = static_cast<H3 *>(p);
    if (h->precision == 4)
      orc3::forward<float>(*static_cast<orc3::Asm<float> *>(h->a), nsrc, src, nrec, rec, nstep, dt, nsb, ntypes,
                           types, seis, t0, final_fields);
    else
      orc3::forward<double>(*static_cast<orc3::Asm<double> *>(h->a), nsrc, src, nrec, rec, nstep, dt, nsb, ntypes,
                            types, seis, t0, final_fields);
    return 0;
  } catch (const std::exception &e) {
    g_err3 = e.what();
    return -1;
  }
}

void orc3d_apply_stiffness(void *p, const double *u, double *out) {
  auto *h = static_cast<H3 *>(p);
  if (h->precision == 4)
    orc3::apply_stiffness<float>(*static_cast<orc3::Asm<float> *>(h->a), u, out);
  else
    orc3::apply_stiffness<double>(*static_cast<orc3::Asm<double> *>(h->a), u, out);
}

}  // extern "C"
