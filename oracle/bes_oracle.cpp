// CPU oracle for the BES residual / implicit-update path.
//
// TEST INFRASTRUCTURE ONLY (see bes_oracle.h).  Each routine restates the
// reference algorithm and cites the reference file:line it follows.  The
// arithmetic is written in the reference's evaluation order so that, compiled
// without FMA contraction, results are bitwise comparable with the reference.
#include "bes_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const int BT_INTERNAL = 1, BT_SLIP = 2, BT_FREESTREAM = 3;

// ---------------------------------------------------------------------------
// Forward-mode AD scalar with the derivative rules of Eigen 3.3.4
// AutoDiffScalar (unsupported/Eigen/src/AutoDiff/AutoDiffScalar.h:181-380,
// 577-588).  Comparisons use values only.
// ---------------------------------------------------------------------------
const int NAD = 12;
struct AD {
  double v;
  double d[NAD];
  AD() : v(0) { for (int i = 0; i < NAD; i++) d[i] = 0; }
  AD(double x) : v(x) { for (int i = 0; i < NAD; i++) d[i] = 0; }  // constant
};
inline AD operator+(const AD &a, const AD &b) { AD r; r.v = a.v + b.v; for (int i = 0; i < NAD; i++) r.d[i] = a.d[i] + b.d[i]; return r; }
inline AD operator-(const AD &a, const AD &b) { AD r; r.v = a.v - b.v; for (int i = 0; i < NAD; i++) r.d[i] = a.d[i] - b.d[i]; return r; }
inline AD operator-(const AD &a) { AD r; r.v = -a.v; for (int i = 0; i < NAD; i++) r.d[i] = -a.d[i]; return r; }
inline AD operator*(const AD &a, const AD &b) { AD r; r.v = a.v * b.v; for (int i = 0; i < NAD; i++) r.d[i] = a.d[i] * b.v + b.d[i] * a.v; return r; }
inline AD operator/(const AD &a, const AD &b) {
  AD r; r.v = a.v / b.v;
  double s = 1.0 / (b.v * b.v);
  for (int i = 0; i < NAD; i++) r.d[i] = (a.d[i] * b.v - b.d[i] * a.v) * s;
  return r;
}
inline AD operator+(const AD &a, double s) { AD r = a; r.v = a.v + s; return r; }
inline AD operator+(double s, const AD &a) { AD r = a; r.v = s + a.v; return r; }
inline AD operator-(const AD &a, double s) { AD r = a; r.v = a.v - s; return r; }
inline AD operator-(double s, const AD &a) { AD r; r.v = s - a.v; for (int i = 0; i < NAD; i++) r.d[i] = -a.d[i]; return r; }
inline AD operator*(const AD &a, double s) { AD r; r.v = a.v * s; for (int i = 0; i < NAD; i++) r.d[i] = a.d[i] * s; return r; }
inline AD operator*(double s, const AD &a) { AD r; r.v = a.v * s; for (int i = 0; i < NAD; i++) r.d[i] = a.d[i] * s; return r; }
inline AD operator/(const AD &a, double s) { AD r; r.v = a.v / s; double q = 1.0 / s; for (int i = 0; i < NAD; i++) r.d[i] = a.d[i] * q; return r; }
inline AD operator/(double s, const AD &a) {
  AD r; r.v = s / a.v;
  double q = (-s) / (a.v * a.v);
  for (int i = 0; i < NAD; i++) r.d[i] = a.d[i] * q;
  return r;
}
inline AD &operator+=(AD &a, const AD &b) { a = a + b; return a; }
inline AD &operator-=(AD &a, const AD &b) { a = a - b; return a; }
inline AD &operator/=(AD &a, double s) { a = a / s; return a; }
inline AD sqrt(const AD &a) {
  AD r; double s = std::sqrt(a.v); r.v = s;
  double q = 0.5 / s;
  for (int i = 0; i < NAD; i++) r.d[i] = a.d[i] * q;
  return r;
}
inline AD abs(const AD &a) {
  AD r; r.v = std::fabs(a.v);
  double sg = a.v < 0 ? -1.0 : 1.0;
  for (int i = 0; i < NAD; i++) r.d[i] = a.d[i] * sg;
  return r;
}
inline bool operator>(const AD &a, const AD &b) { return a.v > b.v; }
inline bool operator>(const AD &a, double b) { return a.v > b; }
inline bool operator>=(const AD &a, double b) { return a.v >= b; }
inline bool operator<=(const AD &a, double b) { return a.v <= b; }
inline bool operator<(const AD &a, double b) { return a.v < b; }
inline double sqrt(double x) { return std::sqrt(x); }
inline double abs(double x) { return std::fabs(x); }

// ---------------------------------------------------------------------------
// Perfect gas, non-dimensional Euler (GasPerfect.tpp, Gas.tpp).
// Conserved [rho, rho u_d, rho E]; primitive [rho, u_d, p, T].
// ---------------------------------------------------------------------------
template <typename T>
struct Gas {
  int nd;
  double gamma, R, cv, cp;
  Gas(int nd_, const or_gas *g) : nd(nd_), gamma(g->gamma), R(g->R) {
    cv = R / (gamma - 1.0);   // GasPerfect.tpp:21-28
    cp = cv + R;
  }
  int iE() const { return nd + 1; }
  int iP() const { return nd + 1; }
  int iT() const { return nd + 2; }
  // GasPerfect.tpp:48-64
  T pressure(const T *U) const {
    T t = 0.0;
    for (int d = 0; d < nd; d++) t += U[1 + d] * U[1 + d];
    return (gamma - 1.0) * (U[iE()] - t / (2.0 * U[0]));
  }
  // GasPerfect.tpp:67-85
  T temperature(const T *U) const {
    T t = 0.0;
    for (int d = 0; d < nd; d++) t += U[1 + d] * U[1 + d];
    return (1.0 / (cv * U[0])) * (U[iE()] - t / (2.0 * U[0]));
  }
  T velocity(const T *U, int d) const { return U[1 + d] / U[0]; }   // GasPerfect.tpp:41-45
  // GasPerfect.tpp:88-111
  void primitive(const T *U, T *P) const {
    P[0] = U[0];
    for (int d = 0; d < nd; d++) P[1 + d] = velocity(U, d);
    P[iP()] = pressure(U);
    P[iT()] = temperature(U);
  }
  T temperatureFromPR(const T &p, const T &rho) const { return p / (R * rho); }   // :114-116
  // GasPerfect.tpp:119-131
  T velocitySquared(const T *U) const {
    T t = 0.0;
    for (int d = 0; d < nd; d++) t += velocity(U, d) * velocity(U, d);
    return t;
  }
  T totalEnthalpy(const T *U) const {   // GasPerfect.tpp:134-146
    return U[iE()] / U[0] + pressure(U) / U[0];
  }
  T soundSpeedFromT(const T &temp) const { return sqrt(gamma * R * temp); }   // :149-151
  T soundSpeed(const T *U) const {   // GasPerfect.tpp:154-160
    return sqrt(gamma * (gamma - 1) * (totalEnthalpy(U) - velocitySquared(U) / 2.0));
  }
  T contravariantFromConserved(const T *U, const double *n) const {   // :163-174
    T t = 0.0;
    for (int d = 0; d < nd; d++) t += velocity(U, d) * n[d];
    return t;
  }
  T machNum(const T &vel, const T &temp) const { return vel / soundSpeedFromT(temp); }   // :177-179
  T contravariantFromPrimitive(const T *P, const double *n) const {   // Gas.tpp:30-41
    T t = 0.0;
    for (int d = 0; d < nd; d++) t += P[1 + d] * n[d];
    return t;
  }
  // GasPerfect.tpp:219-241 (energy), Gas.tpp:44-74 (mass, momentum), :265-291 (all)
  void flux(const T *P, const double *n, T *F) const {
    T V = contravariantFromPrimitive(P, n);
    F[0] = P[0] * V;
    for (int d = 0; d < nd; d++) F[1 + d] = P[0] * P[1 + d] * V + P[iP()] * n[d];
    T q2 = 0.0;
    for (int d = 0; d < nd; d++) q2 += P[1 + d] * P[1 + d];
    F[iE()] = P[0] * (cp * P[iT()] + q2 / 2.0) * V;
  }
};

// ---------------------------------------------------------------------------
// Roe flux-difference splitting with Harten fix (FluxRoe.tpp:61-333).
// Extended primitive layout: [rho, u_d, p, T, H, c, V, q2].
// ---------------------------------------------------------------------------
template <typename T>
void roeFlux(const Gas<T> &g, double harten, const T *UL, const T *UR, const double *n, T *F) {
  const int nd = g.nd, nv = nd + 2, np = nd + 3;
  const int iH = np, iC = np + 1, iV = np + 2, iQ = np + 3, iP = nd + 1;
  T PL[10], PR[10], D[10], A[10], tmp[5];
  // FluxRoe.tpp:61-76
  auto ext = [&](const T *U, T *P) {
    g.primitive(U, P);
    P[iH] = g.totalEnthalpy(U);
    P[iC] = g.soundSpeed(U);
    P[iV] = g.contravariantFromConserved(U, n);
    P[iQ] = g.velocitySquared(U);
  };
  ext(UL, PL);
  ext(UR, PR);
  for (int i = 0; i <= iQ; i++) D[i] = PR[i] - PL[i];   // :315-318
  // Roe averages, FluxRoe.tpp:89-143
  T den = sqrt(PL[0]) + sqrt(PR[0]);
  A[iQ] = 0.0;
  A[iV] = 0.0;
  A[0] = sqrt(PL[0] * PR[0]);
  for (int d = 0; d < nd; d++) {
    A[1 + d] = (PL[1 + d] * sqrt(PL[0]) + PR[1 + d] * sqrt(PR[0])) / den;
    A[iQ] += A[1 + d] * A[1 + d];
    A[iV] += A[1 + d] * n[d];
  }
  A[iH] = (PL[iH] * sqrt(PL[0]) + PR[iH] * sqrt(PR[0])) / den;
  A[iC] = sqrt((T(g.gamma) - 1.0) * (A[iH] - A[iQ] / 2.0));
  // convective part, FluxRoe.tpp:145-155
  for (int i = 0; i < nv; i++) F[i] = 0.0;
  g.flux(PL, n, tmp);
  for (int i = 0; i < nv; i++) F[i] += tmp[i];
  g.flux(PR, n, tmp);
  for (int i = 0; i < nv; i++) F[i] += tmp[i];
  // Harten entropy fix, FluxRoe.tpp:157-171
  auto fix = [&](const T &lam) -> T {
    T delta = abs(A[iC] * harten);
    if (lam > delta) return lam;
    return (lam * lam + delta * delta) / (2.0 * delta);
  };
  // |dF1|, FluxRoe.tpp:173-209
  {
    T c = fix(abs(A[iV] - A[iC])) * (D[iP] - A[0] * A[iC] * D[iV]) / (2.0 * A[iC] * A[iC]);
    F[0] -= c;
    for (int d = 0; d < nd; d++) F[1 + d] -= c * (A[1 + d] - A[iC] * n[d]);
    F[nv - 1] -= c * (A[iH] - A[iC] * A[iV]);
  }
  // |dF234|, FluxRoe.tpp:211-266
  {
    T lam = fix(abs(A[iV]));
    T c1 = lam * (D[0] - D[iP] / (A[iC] * A[iC]));
    T c2 = lam * A[0];
    T t = 0.0;
    F[0] -= c1;
    for (int d = 0; d < nd; d++) {
      F[1 + d] -= (c1 * A[1 + d] + c2 * (D[1 + d] - D[iV] * n[d]));
      t += D[1 + d] * A[1 + d];
    }
    F[nv - 1] -= (c1 * A[iQ] / 2.0 + c2 * (t - D[iV] * A[iV]));
  }
  // |dF5|, FluxRoe.tpp:268-304
  {
    T c = fix(abs(A[iV] + A[iC])) * (D[iP] + A[0] * A[iC] * D[iV]) / (2.0 * A[iC] * A[iC]);
    F[0] -= c;
    for (int d = 0; d < nd; d++) F[1 + d] -= c * (A[1 + d] + A[iC] * n[d]);
    F[nv - 1] -= c * (A[iH] + A[iC] * A[iV]);
  }
  for (int i = 0; i < nv; i++) F[i] /= 2.0;   // :329-331
}

// Slip wall, FluxSlipWall.tpp:32-53
template <typename T>
void slipFlux(const Gas<T> &g, const T *UL, const double *n, T *F) {
  const int nd = g.nd, nv = nd + 2;
  for (int i = 0; i < nv; i++) F[i] = 0.0;
  T p = g.pressure(UL);
  for (int d = 0; d < nd; d++) F[1 + d] = p * n[d];
}

// Characteristic far field, FluxFreestream.tpp:34-226.  Pinf = primitive [nv+1].
template <typename T>
void freestreamFlux(const Gas<T> &g, const T *UL, const T *Pinf, const double *n, T *F) {
  const int nd = g.nd, iP = nd + 1, iT = nd + 2;
  T PL[6], PB[6];
  g.primitive(UL, PL);
  T VL = g.contravariantFromPrimitive(PL, n);
  T ML = g.machNum(abs(VL), PL[iT]);
  if (ML >= 1 && VL <= 0) {           // supersonic inlet :35-39
    g.flux(Pinf, n, F);
  } else if (ML >= 1 && VL > 0) {     // supersonic outlet :42-46
    g.flux(PL, n, F);
  } else if (ML < 1 && VL <= 0) {     // subsonic inlet :49-126
    T cL = g.soundSpeedFromT(PL[iT]);
    T t = 0.0;
    for (int d = 0; d < nd; d++) t += n[d] * (Pinf[1 + d] - PL[1 + d]);
    T c1 = PL[0] * cL;
    PB[iP] = (Pinf[iP] + PL[iP] - c1 * t) / 2.0;
    T c2 = (Pinf[iP] - PB[iP]) / c1;
    PB[0] = Pinf[0] + (PB[iP] - Pinf[iP]) / (cL * cL);
    PB[iT] = g.temperatureFromPR(PB[iP], PB[0]);
    for (int d = 0; d < nd; d++) PB[1 + d] = Pinf[1 + d] - c2 * n[d];
    g.flux(PB, n, F);
  } else if (ML < 1 && VL > 0) {      // subsonic outlet :129-190
    T cL = g.soundSpeedFromT(PL[iT]);
    T c1 = PL[0] * cL;
    PB[iP] = Pinf[iP];
    T c2 = (PL[iP] - PB[iP]) / c1;
    PB[0] = PL[0] + (PB[iP] - PL[iP]) / (cL * cL);
    PB[iT] = g.temperatureFromPR(PB[iP], PB[0]);
    for (int d = 0; d < nd; d++) PB[1 + d] = PL[1 + d] + c2 * n[d];
    g.flux(PB, n, F);
  } else {
    for (int i = 0; i < nd + 2; i++) F[i] = NAN;   // NaN state: flagged downstream
  }
}

template <typename T>
void anyFlux(const Gas<T> &g, double harten, int kind, const T *UL, const T *UR,
             const double *n, T *F) {
  if (kind == BT_INTERNAL) roeFlux(g, harten, UL, UR, n, F);
  else if (kind == BT_SLIP) slipFlux(g, UL, n, F);
  else freestreamFlux(g, UL, UR, n, F);
}

// AD Jacobian of one face flux (FluxJacobianWrapper.tpp:32-82): seeds the
// left inputs as directions 0..nv-1 and, for Roe, the right as nv..2nv-1.
void jacobianOne(int nd, int kind, const double *UL, const double *UR, const double *n,
                 const or_gas *gas, double *JL, double *JR) {
  const int nv = nd + 2;
  Gas<AD> g(nd, gas);
  AD L[5], R[6], F[5];
  for (int i = 0; i < nv; i++) { L[i] = AD(UL[i]); L[i].d[i] = 1.0; }
  int nright = kind == BT_INTERNAL ? nv : (kind == BT_FREESTREAM ? nv + 1 : 0);
  for (int i = 0; i < nright; i++) {
    R[i] = AD(UR[i]);
    R[i].d[nv + i] = 1.0;
  }
  anyFlux(g, gas->harten, kind, L, R, n, F);
  for (int o = 0; o < nv; o++)
    for (int i = 0; i < nv; i++) {
      JL[o * nv + i] = F[o].d[i];
      if (JR) JR[o * nv + i] = kind == BT_INTERNAL ? F[o].d[nv + i] : 0.0;
    }
}

double dist(const double *a, const double *b, int nd) {   // MeshCommon.cpp:21-31
  double s = 0;
  for (int i = 0; i < nd; i++) s += std::pow(a[i] - b[i], 2.0);
  return std::sqrt(s);
}
double magnitude(const double *v, int nd) {               // MeshCommon.cpp:45-52
  double s = 0;
  for (int i = 0; i < nd; i++) s += std::pow(v[i], 2.0);
  return std::sqrt(s);
}
// (v3 - v2) x (v1 - v2), MeshCommon.cpp:77-96
void cross3(double *r, const double *a, const double *b, const double *c) {
  r[0] = (c[1] - b[1]) * (a[2] - b[2]) - (a[1] - b[1]) * (c[2] - b[2]);
  r[1] = (a[0] - b[0]) * (c[2] - b[2]) - (c[0] - b[0]) * (a[2] - b[2]);
  r[2] = (c[0] - b[0]) * (a[1] - b[1]) - (a[0] - b[0]) * (c[1] - b[1]);
}

const int TET_FACES[4][3] = {{0, 2, 1}, {1, 2, 3}, {0, 3, 2}, {0, 1, 3}};   // Mesh.cpp:142-146

}  // namespace

struct or_mesh {
  int nd, nv, k, mode;
  long ncell, nface;
  std::vector<long> owner, nbr;
  std::vector<int> btype;
  std::vector<double> S, n, area, center;
  std::vector<double> vol, centroid, w, r, sproj;
  std::vector<long> cface, cadj, rsel;
  std::vector<int> ownside;   // per (cell, local face): 1 if the cell owns the face
};

extern "C" {

// Mesh::readGmsh (Mesh.cpp:95-558) + completeInitialization_ (:585-636).
or_mesh *or_mesh_create(int nd, long nnode, const double *X, long ncell, const long *cells,
                        long nbface, const long *bfaces, const int *btype, int mode) {
  (void)nnode;
  or_mesh *m = new or_mesh();
  m->nd = nd; m->nv = nd + 2; m->k = nd + 1; m->mode = mode; m->ncell = ncell;
  const int k = m->k;
  std::map<std::vector<long>, long> key2face;
  std::vector<std::vector<long> > fverts;
  auto newFace = [&](const long *v, int nvx, int bt) {
    long id = (long)fverts.size();
    fverts.push_back(std::vector<long>(v, v + nvx));
    m->btype.push_back(bt);
    m->owner.push_back(-1);
    m->nbr.push_back(-1);
    double c[3] = {0, 0, 0}, s[3] = {0, 0, 0};
    if (nd == 2) {   // FaceGmshOne, MeshGmsh.cpp:26-42
      const double *a = X + v[0] * nd, *b = X + v[1] * nd;
      for (int i = 0; i < 2; i++) c[i] = 0.5 * (a[i] + b[i]);
      s[0] = b[1] - a[1];
      s[1] = a[0] - b[0];
    } else {         // FaceGmshTwo, MeshGmsh.cpp:49-68
      const double *a = X + v[0] * 3, *b = X + v[1] * 3, *cc = X + v[2] * 3;
      for (int i = 0; i < 3; i++) c[i] = (a[i] + b[i] + cc[i]) / 3;
      cross3(s, a, b, cc);
      for (int i = 0; i < 3; i++) s[i] /= 2.0;
    }
    double ar = magnitude(s, nd);
    for (int i = 0; i < nd; i++) {
      m->S.push_back(s[i]);
      m->n.push_back(s[i] / ar);
      m->center.push_back(c[i]);
    }
    m->area.push_back(ar);
    return id;
  };
  auto key = [](const long *v, int nvx) {
    std::vector<long> kk(v, v + nvx);
    std::sort(kk.begin(), kk.end());
    return kk;
  };
  for (long b = 0; b < nbface; b++) {   // boundary elements, file order
    const long *v = bfaces + b * nd;
    long id = newFace(v, nd, btype[b]);
    key2face[key(v, nd)] = id;
  }
  m->cface.assign(ncell * k, -1);
  for (long e = 0; e < ncell; e++) {   // volume elements, file order
    const long *cv = cells + e * k;
    for (int mm = 0; mm < k; mm++) {
      long fv[3];
      if (nd == 2) { fv[0] = cv[mm]; fv[1] = cv[(mm + 1) % 3]; }
      else for (int q = 0; q < 3; q++) fv[q] = cv[TET_FACES[mm][q]];
      std::vector<long> kk = key(fv, nd);
      auto it = key2face.find(kk);
      long f;
      if (it != key2face.end()) {
        f = it->second;
        double out[3];
        if (nd == 2) {   // MeshCommon.cpp:60-67
          const double *a = X + fv[0] * 2, *b = X + fv[1] * 2;
          out[0] = b[1] - a[1];
          out[1] = -b[0] + a[0];
        } else {
          cross3(out, X + fv[0] * 3, X + fv[1] * 3, X + fv[2] * 3);
        }
        if (m->btype[f] != BT_INTERNAL) {   // Mesh.cpp:365-378
          double dp = 0;
          for (int i = 0; i < nd; i++) dp += m->S[f * nd + i] * out[i];
          if (dp < 0)
            for (int i = 0; i < nd; i++) { m->S[f * nd + i] = -m->S[f * nd + i]; m->n[f * nd + i] = -m->n[f * nd + i]; }
          m->owner[f] = e;
          m->nbr[f] = -1;
        } else {
          m->nbr[f] = e;
        }
      } else {
        f = newFace(fv, nd, BT_INTERNAL);
        key2face[kk] = f;
        m->owner[f] = e;
      }
      m->cface[e * k + mm] = f;
    }
  }
  m->nface = (long)fverts.size();
  // completeInitialization_, Mesh.cpp:585-636
  m->vol.resize(ncell);
  m->centroid.resize(ncell * nd);
  m->r.resize(ncell * k * nd);
  m->cadj.assign(ncell * k, -1);
  m->w.assign(ncell * k, 0.0);
  m->rsel.resize(ncell * k);
  m->sproj.assign(ncell * nd, 0.0);
  m->ownside.resize(ncell * k);
  for (long e = 0; e < ncell; e++) {
    const long *cv = cells + e * k;
    double *cen = &m->centroid[e * nd];
    if (nd == 2) {   // MeshGmsh.cpp:76-93
      const double *a = X + cv[0] * 2, *b = X + cv[1] * 2, *c = X + cv[2] * 2;
      for (int i = 0; i < 2; i++) cen[i] = (a[i] + b[i] + c[i]) / 3;
      m->vol[e] = 0.5 * std::fabs((a[0] - b[0]) * (c[1] - b[1]) - (c[0] - b[0]) * (a[1] - b[1]));
    } else {         // MeshGmsh.cpp:101-145
      const double *p[4] = {X + cv[0] * 3, X + cv[1] * 3, X + cv[2] * 3, X + cv[3] * 3};
      for (int i = 0; i < 3; i++) {
        if (mode & 2) cen[i] = (p[0][i] + p[1][i] + p[2][i] + p[3][i]) / 4;
        else cen[i] = (p[0][i] + p[1][i] + p[2][i] + p[2][i]) / 4;
      }
      double vol = 0;
      for (int f = 0; f < 4; f++) {
        double s[3], fc[3];
        cross3(s, p[TET_FACES[f][0]], p[TET_FACES[f][1]], p[TET_FACES[f][2]]);
        for (int i = 0; i < 3; i++) {
          s[i] /= 2.0;
          fc[i] = 0;
          for (int q = 0; q < 3; q++) fc[i] += p[TET_FACES[f][q]][i];
          fc[i] /= 3.0;
        }
        double dp = 0;
        for (int i = 0; i < 3; i++) dp += fc[i] * s[i];
        vol += dp;
      }
      m->vol[e] = (mode & 1) ? vol / 3.0 : vol;
    }
    for (int mm = 0; mm < k; mm++) {   // Element.cpp:111-118
      long f = m->cface[e * k + mm];
      for (int i = 0; i < nd; i++) m->r[(e * k + mm) * nd + i] = m->center[f * nd + i] - cen[i];
    }
  }
  for (long e = 0; e < ncell; e++) {
    for (int mm = 0; mm < k; mm++) {
      long f = m->cface[e * k + mm];
      m->ownside[e * k + mm] = (m->owner[f] == e) ? 1 : 0;
      if (m->btype[f] == BT_INTERNAL)
        m->cadj[e * k + mm] = (m->owner[f] != e) ? m->owner[f] : m->nbr[f];
      long j = m->cadj[e * k + mm];
      if (j >= 0) {   // Element.cpp:120-140
        double dO = dist(&m->centroid[e * nd], &m->center[f * nd], nd);
        double dN = dist(&m->centroid[j * nd], &m->center[f * nd], nd);
        m->w[e * k + mm] = dN / (dO + dN);
      }
      // reconstruction vector lookup (Element.cpp:102-105,181-187): the face map
      // is keyed before face ids exist, so every lookup falls back to local
      // face 0 except global face 0, which hits the last local face.
      if (mode & 4) m->rsel[e * k + mm] = mm;
      else m->rsel[e * k + mm] = (f == 0) ? k - 1 : 0;
      for (int i = 0; i < nd; i++)   // TimeStepCalculatorLocal.cpp:51-65
        m->sproj[e * nd + i] += std::fabs(m->S[f * nd + i]);
    }
    for (int i = 0; i < nd; i++) m->sproj[e * nd + i] /= 2.0;
  }
  return m;
}

void or_mesh_destroy(or_mesh *m) { delete m; }
long or_mesh_nface(const or_mesh *m) { return m->nface; }
int or_mesh_kface(const or_mesh *m) { return m->k; }

void or_mesh_get_faces(const or_mesh *m, long *owner, long *nbr, int *btype, double *S,
                       double *n, double *area, double *center) {
  if (owner) std::copy(m->owner.begin(), m->owner.end(), owner);
  if (nbr) std::copy(m->nbr.begin(), m->nbr.end(), nbr);
  if (btype) std::copy(m->btype.begin(), m->btype.end(), btype);
  if (S) std::copy(m->S.begin(), m->S.end(), S);
  if (n) std::copy(m->n.begin(), m->n.end(), n);
  if (area) std::copy(m->area.begin(), m->area.end(), area);
  if (center) std::copy(m->center.begin(), m->center.end(), center);
}

void or_mesh_get_cells(const or_mesh *m, double *vol, double *centroid, long *cface, long *cadj,
                       double *w, double *r, long *rsel, double *sproj) {
  if (vol) std::copy(m->vol.begin(), m->vol.end(), vol);
  if (centroid) std::copy(m->centroid.begin(), m->centroid.end(), centroid);
  if (cface) std::copy(m->cface.begin(), m->cface.end(), cface);
  if (cadj) std::copy(m->cadj.begin(), m->cadj.end(), cadj);
  if (w) std::copy(m->w.begin(), m->w.end(), w);
  if (r) std::copy(m->r.begin(), m->r.end(), r);
  if (rsel) std::copy(m->rsel.begin(), m->rsel.end(), rsel);
  if (sproj) std::copy(m->sproj.begin(), m->sproj.end(), sproj);
}

void or_flux(int nd, int kind, long n, const double *UL, const double *UR, const double *nrm,
             const or_gas *gas, double *F) {
  const int nv = nd + 2;
  Gas<double> g(nd, gas);
  for (long i = 0; i < n; i++)
    anyFlux(g, gas->harten, kind, UL + i * nv, UR + i * (nv + 1), nrm + i * nd, F + i * nv);
}

void or_jacobian(int nd, int kind, long n, const double *UL, const double *UR, const double *nrm,
                 const or_gas *gas, double *JL, double *JR) {
  const int nv = nd + 2;
  for (long i = 0; i < n; i++)
    jacobianOne(nd, kind, UL + i * nv, UR + i * (nv + 1), nrm + i * nd, gas, JL + i * nv * nv,
                JR ? JR + i * nv * nv : nullptr);
}

// GradientGreenGauss.cpp:21-135 (cell loop)
void or_gradient(const or_mesh *m, const double *U, double *G) {
  const int nd = m->nd, nv = m->nv, k = m->k;
#pragma omp parallel for
  for (long e = 0; e < m->ncell; e++) {
    double *g = G + e * nv * nd;
    for (int q = 0; q < nv * nd; q++) g[q] = 0;
    for (int mm = 0; mm < k; mm++) {
      long f = m->cface[e * k + mm];
      const double *S = &m->S[f * nd];
      long j = m->cadj[e * k + mm];
      double w = m->w[e * k + mm];
      for (int v = 0; v < nv; v++)
        for (int d = 0; d < nd; d++) {
          if (j < 0) g[v * nd + d] += S[d] * (U[e * nv + v]);
          else if (m->owner[f] == e) g[v * nd + d] += S[d] * (w * U[e * nv + v] + (1 - w) * U[j * nv + v]);
          else g[v * nd + d] -= S[d] * (w * U[e * nv + v] + (1 - w) * U[j * nv + v]);
        }
      if (m->btype[f] == BT_SLIP) {   // tangential-velocity correction :82-121
        const double *n = &m->n[f * nd];
        double vn = 0, t[3];
        for (int d = 0; d < nd; d++) vn += U[e * nv + 1 + d] * n[d];
        for (int d = 0; d < nd; d++) t[d] = U[e * nv + 1 + d] - vn * n[d];
        for (int a = 0; a < nd; a++)
          for (int d = 0; d < nd; d++)
            g[(1 + a) * nd + d] += (t[a] * S[d] - U[e * nv + 1 + a] * S[d]);
      }
    }
    double V = m->vol[e];
    for (int q = 0; q < nv * nd; q++) g[q] /= V;
  }
}

// LimiterVenkatakrishnan.cpp:21-128
static double faceLimiter(double mn, double mx, double d2, double eps) {
  if (d2 > 0)
    return (1 / d2) * ((mx * mx + eps * eps) * d2 + 2.0 * d2 * d2 * mx) /
           (mx * mx + 2.0 * d2 * d2 + mx * d2 + eps * eps);
  else if (d2 < 0)
    return (1 / d2) * ((mn * mn + eps * eps) * d2 + 2.0 * d2 * d2 * mn) /
           (mn * mn + 2.0 * d2 * d2 + mn * d2 + eps * eps);
  return 1.0;
}

void or_limiter(const or_mesh *m, const double *U, const double *G, const or_gas *gas, double *phi) {
  const int nd = m->nd, nv = m->nv, k = m->k;
#pragma omp parallel for
  for (long e = 0; e < m->ncell; e++) {
    double L = std::pow(m->vol[e], 1.0 / nd);
    double eps = std::pow(gas->venkat_k * L, 3.0 / 2.0);
    for (int v = 0; v < nv; v++) {
      double u = U[e * nv + v], mn = u, mx = u;
      for (int mm = 0; mm < k; mm++) {
        long j = m->cadj[e * k + mm];
        if (j >= 0) {
          double x = U[j * nv + v];
          mn = (x < mn) ? x : mn;
          mx = (x > mx) ? x : mx;
        }
      }
      mn -= u;
      mx -= u;
      double best = 1.0;
      for (int mm = 0; mm < k; mm++) {
        if (m->cadj[e * k + mm] < 0) continue;
        double d2 = 0;
        for (int d = 0; d < nd; d++) d2 += G[(e * nv + v) * nd + d] * m->r[(e * k + mm) * nd + d];
        d2 /= 2.0;
        double f = faceLimiter(mn, mx, d2, eps);
        if (f < best) best = f;
      }
      phi[e * nv + v] = best;
    }
  }
}

// TimeStepCalculatorLocal.cpp:67-93, TimeStepCalculatorGlobal.cpp:64-98
void or_timestep(const or_mesh *m, const double *U, double cfl, int local, const or_gas *gas,
                 double *dt) {
  const int nd = m->nd, nv = m->nv;
  Gas<double> g(nd, gas);
  double gmin = 0;
  for (long e = 0; e < m->ncell; e++) {
    const double *u = U + e * nv;
    double sr = 0.0, c = g.soundSpeed(u);
    for (int d = 0; d < nd; d++) {
      double vel = std::fabs(g.velocity(u, d));
      sr += (c + vel) * m->sproj[e * nd + d];
    }
    double t = cfl * m->vol[e] / sr;
    dt[e] = t;
    if (e == 0) gmin = t;
    else gmin = (t < gmin) ? t : gmin;
  }
  if (!local)
    for (long e = 0; e < m->ncell; e++) dt[e] = gmin;
}

} // extern "C"

namespace {
// Face states, Solver.cpp:69-121 (second order) with the reconstruction
// vector chosen by rsel.
void faceStates(const or_mesh *m, const double *U, const double *G, const double *phi, long f,
                double *L, double *R) {
  const int nd = m->nd, nv = m->nv, k = m->k;
  long sides[2] = {m->owner[f], m->nbr[f]};
  double *out[2] = {L, R};
  for (int s = 0; s < 2; s++) {
    long e = sides[s];
    if (e < 0) { for (int v = 0; v < nv; v++) out[s][v] = 0; continue; }
    int loc = 0;
    for (int mm = 0; mm < k; mm++) if (m->cface[e * k + mm] == f) loc = mm;
    const double *r = &m->r[(e * k + m->rsel[e * k + loc]) * nd];
    for (int v = 0; v < nv; v++) {
      double p = phi[e * nv + v];
      out[s][v] = U[e * nv + v];
      for (int d = 0; d < nd; d++) out[s][v] += p * G[(e * nv + v) * nd + d] * r[d];
    }
  }
}
}  // namespace

extern "C" {

// Face loop of SolverExplicit.cpp:285-344 (serial, face order).
void or_residual(const or_mesh *m, const double *U, const double *G, const double *phi,
                 const double *finf, const or_gas *gas, double *R, double *F) {
  const int nd = m->nd, nv = m->nv;
  Gas<double> g(nd, gas);
  std::fill(R, R + m->ncell * nv, 0.0);
  double L[5], Rv[5], fl[5];
  for (long f = 0; f < m->nface; f++) {
    faceStates(m, U, G, phi, f, L, Rv);
    anyFlux(g, gas->harten, m->btype[f], L, m->btype[f] == BT_FREESTREAM ? finf : Rv,
            &m->n[f * nd], fl);
    double A = m->area[f];
    long o = m->owner[f], nb = m->nbr[f];
    for (int v = 0; v < nv; v++) {
      if (F) F[f * nv + v] = fl[v];
      R[o * nv + v] += fl[v] * A;
      if (nb >= 0) R[nb * nv + v] -= fl[v] * A;
    }
  }
}

// SolverExplicit::timeStep_, SolverExplicit.cpp:259-393.  The face loop is
// face-parallel with atomic scatters, as in the reference.
void or_explicit(const or_mesh *m, double *U, int niter, int nstage, const double *alpha,
                 double cfl, int local, const double *finf, const or_gas *gas, double *linf,
                 int nthreads) {
  const int nd = m->nd, nv = m->nv;
  const long ne = m->ncell;
#ifdef _OPENMP
  int prev = omp_get_max_threads();
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  std::vector<double> U1(U, U + ne * nv), G(ne * nv * nd), P(ne * nv), sub(ne * nv), dt(ne);
  Gas<double> g(nd, gas);
  for (int it = 0; it < niter; it++) {
    or_timestep(m, U, cfl, local, gas, dt.data());
    for (int s = 0; s < nstage; s++) {
      const double *cur = (s == 0) ? U : U1.data();
      std::fill(sub.begin(), sub.end(), 0.0);
      or_gradient(m, cur, G.data());
      or_limiter(m, cur, G.data(), gas, P.data());
#pragma omp parallel for
      for (long f = 0; f < m->nface; f++) {
        double L[5], Rv[5], fl[5];
        faceStates(m, cur, G.data(), P.data(), f, L, Rv);
        anyFlux(g, gas->harten, m->btype[f], L, m->btype[f] == BT_FREESTREAM ? finf : Rv,
                &m->n[f * nd], fl);
        double A = m->area[f];
        long o = m->owner[f], nb = m->nbr[f];
        for (int v = 0; v < nv; v++) {
#pragma omp atomic
          sub[o * nv + v] += fl[v] * A;
          if (nb >= 0) {
#pragma omp atomic
            sub[nb * nv + v] -= fl[v] * A;
          }
        }
      }
      bool last = s == nstage - 1;
      if (last && linf)
        for (int v = 0; v < nv; v++) {
          double mx = 0;
          for (long e = 0; e < ne; e++) mx = std::fabs(sub[e * nv + v]) > mx ? std::fabs(sub[e * nv + v]) : mx;
          linf[it * nv + v] = mx;
        }
#pragma omp parallel for
      for (long e = 0; e < ne; e++)
        for (int v = 0; v < nv; v++)
          U1[e * nv + v] = U[e * nv + v] - (alpha[s] * dt[e] / m->vol[e]) * sub[e * nv + v];
    }
    std::copy(U1.begin(), U1.end(), U);
  }
#ifdef _OPENMP
  omp_set_num_threads(prev);
#endif
}

long or_implicit_nnzb(const or_mesh *m) {
  long n = 0;
  for (long e = 0; e < m->ncell; e++) {
    n++;
    for (int mm = 0; mm < m->k; mm++) if (m->cadj[e * m->k + mm] >= 0) n++;
  }
  return n;
}

// SolverImplicit.cpp:347-356 (reset), :358-507 (face loop), CoefficientMatrix.cpp:114-204.
void or_implicit_assemble(const or_mesh *m, const double *U, double cfl, int local,
                          const double *finf, const or_gas *gas, int repeat_nv, long *rowptr,
                          long *bcol, double *bval, double *rhs, double *dt_out) {
  const int nd = m->nd, nv = m->nv, k = m->k, bs = nv * nv;
  const long ne = m->ncell;
  std::vector<double> G(ne * nv * nd), P(ne * nv), dt(ne);
  or_timestep(m, U, cfl, local, gas, dt.data());
  or_gradient(m, U, G.data());
  or_limiter(m, U, G.data(), gas, P.data());
  if (dt_out) std::copy(dt.begin(), dt.end(), dt_out);
  // pattern: own block, then face neighbours in local face order
  long nz = 0;
  std::vector<long> slot(ne * k, -1);
  for (long e = 0; e < ne; e++) {
    rowptr[e] = nz;
    bcol[nz++] = e;
    for (int mm = 0; mm < k; mm++) {
      long j = m->cadj[e * k + mm];
      if (j >= 0) { slot[e * k + mm] = nz; bcol[nz++] = j; }
    }
  }
  rowptr[ne] = nz;
  std::fill(bval, bval + nz * bs, 0.0);
  for (long e = 0; e < ne; e++)
    for (int a = 0; a < nv; a++) bval[rowptr[e] * bs + a * nv + a] += 1.0;
  std::fill(rhs, rhs + ne * nv, 0.0);
  auto blockOf = [&](long row, long col) -> long {
    if (row == col) return rowptr[row];
    for (int mm = 0; mm < k; mm++) if (m->cadj[row * k + mm] == col) return slot[row * k + mm];
    return -1;
  };
  const int reps = repeat_nv ? nv : 1;
  auto add = [&](long row, long col, double mul, const double *J) {
    double *B = bval + blockOf(row, col) * bs;
    for (int r = 0; r < reps; r++)
      for (int q = 0; q < bs; q++) B[q] += mul * J[q];
  };
  Gas<double> g(nd, gas);
  double L[5], Rv[5], fl[5], JL[25], JR[25];
  for (long f = 0; f < m->nface; f++) {
    int bt = m->btype[f];
    faceStates(m, U, G.data(), P.data(), f, L, Rv);
    anyFlux(g, gas->harten, bt, L, bt == BT_FREESTREAM ? finf : Rv, &m->n[f * nd], fl);
    long o = m->owner[f], nb = m->nbr[f];
    for (int v = 0; v < nv; v++) {   // Solver.cpp:41-67 (first-order states)
      L[v] = U[o * nv + v];
      Rv[v] = nb >= 0 ? U[nb * nv + v] : 0.0;
    }
    double URb[6];
    const double *ur = Rv;
    if (bt == BT_FREESTREAM) ur = finf;
    for (int v = 0; v < nv + 1; v++) URb[v] = (bt == BT_INTERNAL && v < nv) ? ur[v] : (bt == BT_FREESTREAM ? ur[v] : 0.0);
    jacobianOne(nd, bt, L, URb, &m->n[f * nd], gas, JL, JR);
    double A = m->area[f];
    double mL = dt[o] / m->vol[o] * A;
    for (int v = 0; v < nv; v++) rhs[o * nv + v] -= mL * fl[v];
    add(o, o, mL, JL);
    if (nb >= 0) {
      double mR = dt[nb] / m->vol[nb] * A;
      for (int v = 0; v < nv; v++) rhs[nb * nv + v] += dt[nb] / m->vol[nb] * A * fl[v];
      add(o, nb, mL, JR);
      add(nb, o, -mR, JL);
      add(nb, nb, -mR, JR);
    }
  }
}

} // extern "C"

// ---------------------------------------------------------------------------
// Implicit pseudo-time loop on the CPU (SolverImplicit::solve,
// SolverImplicit.cpp:552-600), with OpenMP where the reference has it.  Used
// as the checker's converged solutions and as the timed CPU baseline.
//   assembly   SolverImplicit.cpp:358-507: face-parallel loop, atomic adds to
//              the right-hand side and to every block entry (MatrixSparse's
//              addToVal is atomic), each block added nV times when repeat_nv
//   solve      the reference calls AMGCL (CoefficientMatrix.cpp:210-325),
//              which needs Boost and cannot be built here; this restatement
//              solves with right-preconditioned GMRES(m), classical
//              Gram-Schmidt, preconditioned by block ILU(0) in a multicolour
//              order (rows of one colour are independent, so each colour is one
//              parallel loop), ||b - A x|| <= rtol ||b|| checked on the true
//              residual before stopping (AMGCL's relative tolerance, Q6)
//   update     SolverImplicit.cpp:509-528 (U += dU, Linf of b)
//   CFL        SolverImplicit.cpp:530-550 (adaptive Courant number)
// ---------------------------------------------------------------------------
struct or_implicit {
  int nv = 0, k = 0, restart = 30, maxiter = 300, nthreads = 0;
  long ne = 0, nnzb = 0;
  double rtol = 1e-3;
  std::vector<long> rowptr, bcol, slot;     // BSR: own block, then neighbours in local face order
  std::vector<int> color;                   // greedy colouring in cell order
  std::vector<std::vector<long>> rows;      // rows of each colour, ascending
  std::vector<double> A, LU, Dinv, b, R, x, dt;
  std::vector<double> V, Zv, w, u;          // Krylov basis [(m+1)][n], work vectors
  std::vector<double> phi_frozen;           // [ne][nv] limiter used instead of recomputing it (if set)
  long last_iters = 0;
};

namespace {

typedef std::vector<double> dvec;

void blk_mul(int nv, const double *A, const double *B, double *C) {   // C = A B
  for (int r = 0; r < nv; r++)
    for (int c = 0; c < nv; c++) {
      double s = 0;
      for (int p = 0; p < nv; p++) s += A[r * nv + p] * B[p * nv + c];
      C[r * nv + c] = s;
    }
}

void blk_inverse(int nv, const double *Ain, double *Inv) {   // Gauss-Jordan, partial pivoting
  double A[25];
  std::copy(Ain, Ain + nv * nv, A);
  for (int r = 0; r < nv; r++)
    for (int c = 0; c < nv; c++) Inv[r * nv + c] = r == c ? 1.0 : 0.0;
  for (int p = 0; p < nv; p++) {
    int piv = p;
    for (int r = p + 1; r < nv; r++)
      if (std::fabs(A[r * nv + p]) > std::fabs(A[piv * nv + p])) piv = r;
    if (piv != p)
      for (int c = 0; c < nv; c++) {
        std::swap(A[p * nv + c], A[piv * nv + c]);
        std::swap(Inv[p * nv + c], Inv[piv * nv + c]);
      }
    const double iv = 1.0 / A[p * nv + p];
    for (int c = 0; c < nv; c++) { A[p * nv + c] *= iv; Inv[p * nv + c] *= iv; }
    for (int r = 0; r < nv; r++) {
      if (r == p) continue;
      const double f = A[r * nv + p];
      for (int c = 0; c < nv; c++) { A[r * nv + c] -= f * A[p * nv + c]; Inv[r * nv + c] -= f * Inv[p * nv + c]; }
    }
  }
}

inline void atomic_add(double *p, double v) {
#pragma omp atomic
  *p += v;
}

// y = A x (BSR rows, cell order)
void imp_spmv(const or_mesh *m, const or_implicit *s, const double *x, double *y) {
  const int nv = s->nv, nb = nv * nv;
#pragma omp parallel for
  for (long e = 0; e < s->ne; e++) {
    double acc[5] = {0, 0, 0, 0, 0};
    for (long q = s->rowptr[e]; q < s->rowptr[e + 1]; q++) {
      const double *B = &s->A[q * nb], *xj = x + s->bcol[q] * nv;
      for (int r = 0; r < nv; r++)
        for (int c = 0; c < nv; c++) acc[r] += B[r * nv + c] * xj[c];
    }
    for (int r = 0; r < nv; r++) y[e * nv + r] = acc[r];
  }
  (void)m;
}

// block ILU(0): rows eliminated colour by colour; a row's lower neighbours in
// the order (colour, cell index); LU holds L_ik in lower slots, U_ij in upper
// slots, Dinv the inverse of each pivot block
void imp_factor(const or_mesh *m, or_implicit *s) {
  const int nv = s->nv, nb = nv * nv, k = s->k;
  s->LU = s->A;
  for (const auto &rows : s->rows) {
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)rows.size(); q++) {
      const long i = rows[q];
      const int ci = s->color[i];
      int ord[4], nl = 0;
      for (int mm = 0; mm < k; mm++) {
        const long j = m->cadj[i * k + mm];
        if (j >= 0 && s->color[j] < ci) ord[nl++] = mm;
      }
      std::sort(ord, ord + nl, [&](int a, int b) {
        const long ja = m->cadj[i * k + a], jb = m->cadj[i * k + b];
        return s->color[ja] != s->color[jb] ? s->color[ja] < s->color[jb] : ja < jb;
      });
      double *Aii = &s->LU[s->rowptr[i] * nb];
      for (int t = 0; t < nl; t++) {
        const int mm = ord[t];
        const long kk = m->cadj[i * k + mm];
        double *Lik = &s->LU[s->slot[i * k + mm] * nb];
        double tmp[25];
        blk_mul(nv, Lik, &s->Dinv[kk * nb], tmp);
        std::copy(tmp, tmp + nb, Lik);
        for (int mk = 0; mk < k; mk++) {
          const long j = m->cadj[kk * k + mk];
          if (j < 0 || s->color[j] < s->color[kk]) continue;   // upper part of row kk only
          double *dst = nullptr;
          if (j == i) dst = Aii;
          else
            for (int m2 = 0; m2 < k; m2++)
              if (m->cadj[i * k + m2] == j) dst = &s->LU[s->slot[i * k + m2] * nb];
          if (!dst) continue;
          double upd[25];
          blk_mul(nv, Lik, &s->LU[s->slot[kk * k + mk] * nb], upd);
          for (int e = 0; e < nb; e++) dst[e] -= upd[e];
        }
      }
      blk_inverse(nv, Aii, &s->Dinv[i * nb]);
    }
  }
}

// z = (LU)^-1 r
void imp_precond(const or_mesh *m, const or_implicit *s, const double *r, double *z, double *y) {
  const int nv = s->nv, nb = nv * nv, k = s->k;
  for (size_t c = 0; c < s->rows.size(); c++) {
    const auto &rows = s->rows[c];
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)rows.size(); q++) {
      const long i = rows[q];
      double acc[5];
      for (int a = 0; a < nv; a++) acc[a] = r[i * nv + a];
      for (int mm = 0; mm < k; mm++) {
        const long j = m->cadj[i * k + mm];
        if (j < 0 || s->color[j] >= (int)c) continue;
        const double *L = &s->LU[s->slot[i * k + mm] * nb];
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++) acc[a] -= L[a * nv + b] * y[j * nv + b];
      }
      for (int a = 0; a < nv; a++) y[i * nv + a] = acc[a];
    }
  }
  for (long c = (long)s->rows.size() - 1; c >= 0; c--) {
    const auto &rows = s->rows[c];
#pragma omp parallel for schedule(static)
    for (long q = 0; q < (long)rows.size(); q++) {
      const long i = rows[q];
      double acc[5];
      for (int a = 0; a < nv; a++) acc[a] = y[i * nv + a];
      for (int mm = 0; mm < k; mm++) {
        const long j = m->cadj[i * k + mm];
        if (j < 0 || s->color[j] <= (int)c) continue;
        const double *U = &s->LU[s->slot[i * k + mm] * nb];
        for (int a = 0; a < nv; a++)
          for (int b = 0; b < nv; b++) acc[a] -= U[a * nv + b] * z[j * nv + b];
      }
      const double *D = &s->Dinv[i * nb];
      for (int a = 0; a < nv; a++) {
        double t = 0;
        for (int b = 0; b < nv; b++) t += D[a * nv + b] * acc[b];
        z[i * nv + a] = t;
      }
    }
  }
}

double imp_dot(const double *a, const double *b, long n) {
  double s = 0;
#pragma omp parallel for reduction(+ : s)
  for (long e = 0; e < n; e++) s += a[e] * b[e];
  return s;
}

// right-preconditioned GMRES(m); x = 0 on entry (SolverImplicit.cpp:349);
// returns iterations, relres = true ||b - A x|| / ||b||
long imp_gmres(const or_mesh *m, or_implicit *s, double *relres) {
  const long n = s->ne * s->nv;
  const int M = s->restart;
  dvec &x = s->x;
  std::fill(x.begin(), x.end(), 0.0);
  double *V = s->V.data(), *z = s->Zv.data(), *y = s->Zv.data() + n, *w = s->w.data(), *u = s->u.data();
  const double bnorm = std::sqrt(imp_dot(s->b.data(), s->b.data(), n));
  *relres = 0;
  if (bnorm == 0) return 0;
  const double target = s->rtol * bnorm;
  std::vector<double> H((size_t)(M + 1) * M), cs(M), sn(M), g(M + 1), yy(M), h(M + 1);
  long total = 0;
  bool first = true;
  double rnorm = bnorm;
  while (total < s->maxiter) {
    if (first) {
#pragma omp parallel for
      for (long e = 0; e < n; e++) V[e] = s->b[e] / bnorm;
      first = false;
    } else {   // true residual of the current x: stop, or restart from it
      imp_spmv(m, s, x.data(), w);
#pragma omp parallel for
      for (long e = 0; e < n; e++) w[e] = s->b[e] - w[e];
      rnorm = std::sqrt(imp_dot(w, w, n));
      if (rnorm <= target) break;
#pragma omp parallel for
      for (long e = 0; e < n; e++) V[e] = w[e] / rnorm;
    }
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = rnorm;
    int j = 0;
    for (; j < M && total < s->maxiter; j++) {
      double *vj = V + (size_t)j * n, *vn = V + (size_t)(j + 1) * n;
      imp_precond(m, s, vj, z, y);
      imp_spmv(m, s, z, vn);
      for (int q = 0; q <= j; q++) h[q] = imp_dot(V + (size_t)q * n, vn, n);   // classical Gram-Schmidt
#pragma omp parallel for
      for (long e = 0; e < n; e++) {
        double t = vn[e];
        for (int q = 0; q <= j; q++) t -= h[q] * V[(size_t)q * n + e];
        vn[e] = t;
      }
      const double hn = std::sqrt(imp_dot(vn, vn, n));
      h[j + 1] = hn;
      if (hn > 0) {
#pragma omp parallel for
        for (long e = 0; e < n; e++) vn[e] /= hn;
      }
      for (int q = 0; q < j; q++) {
        const double t = cs[q] * h[q] + sn[q] * h[q + 1];
        h[q + 1] = -sn[q] * h[q] + cs[q] * h[q + 1];
        h[q] = t;
      }
      const double den = std::hypot(h[j], h[j + 1]);
      cs[j] = den > 0 ? h[j] / den : 1.0;
      sn[j] = den > 0 ? h[j + 1] / den : 0.0;
      h[j] = den;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      for (int q = 0; q <= j; q++) H[(size_t)q * M + j] = h[q];
      total++;
      if (std::fabs(g[j + 1]) <= target || hn == 0) { j++; break; }
    }
    for (int q = j - 1; q >= 0; q--) {
      double t = g[q];
      for (int p = q + 1; p < j; p++) t -= H[(size_t)q * M + p] * yy[p];
      yy[q] = t / H[(size_t)q * M + q];
    }
#pragma omp parallel for
    for (long e = 0; e < n; e++) {
      double t = 0;
      for (int q = 0; q < j; q++) t += yy[q] * V[(size_t)q * n + e];
      u[e] = t;
    }
    imp_precond(m, s, u, z, y);
#pragma omp parallel for
    for (long e = 0; e < n; e++) x[e] += z[e];
  }
  imp_spmv(m, s, x.data(), w);
  double r2 = 0;
#pragma omp parallel for reduction(+ : r2)
  for (long e = 0; e < n; e++) {
    const double t = s->b[e] - w[e];
    r2 += t * t;
  }
  *relres = std::sqrt(r2) / bnorm;
  return total;
}

// SolverImplicit.cpp:347-507 with the face loop in parallel (atomic adds)
void imp_assemble(const or_mesh *m, or_implicit *s, const double *U, double cfl, int local,
                  const double *finf, const or_gas *gas, int repeat_nv) {
  const int nd = m->nd, nv = m->nv, k = m->k, nb = nv * nv;
  const long ne = m->ncell;
  std::vector<double> G(ne * nv * nd), P(ne * nv);
  or_timestep(m, U, cfl, local, gas, s->dt.data());   // serial, as TimeStepCalculatorLocal
  or_gradient(m, U, G.data());
  if (s->phi_frozen.empty()) or_limiter(m, U, G.data(), gas, P.data());
  else P = s->phi_frozen;
  std::fill(s->A.begin(), s->A.end(), 0.0);
  for (long e = 0; e < ne; e++)
    for (int a = 0; a < nv; a++) s->A[s->rowptr[e] * nb + a * nv + a] = 1.0;
  std::fill(s->b.begin(), s->b.end(), 0.0);
  std::fill(s->R.begin(), s->R.end(), 0.0);
  const int reps = repeat_nv ? nv : 1;
  const double *dt = s->dt.data();
  auto blockOf = [&](long row, long col) -> long {
    if (row == col) return s->rowptr[row];
    for (int mm = 0; mm < k; mm++) if (m->cadj[row * k + mm] == col) return s->slot[row * k + mm];
    return -1;
  };
  auto add = [&](long row, long col, double mul, const double *J) {
    double *B = &s->A[blockOf(row, col) * nb];
    for (int r = 0; r < reps; r++)
      for (int q = 0; q < nb; q++) atomic_add(&B[q], mul * J[q]);
  };
  Gas<double> g(nd, gas);
#pragma omp parallel for schedule(dynamic, 256)
  for (long f = 0; f < m->nface; f++) {
    double L[5], Rv[5], fl[5], JL[25], JR[25], URb[6];
    const int bt = m->btype[f];
    faceStates(m, U, G.data(), P.data(), f, L, Rv);
    anyFlux(g, gas->harten, bt, L, bt == BT_FREESTREAM ? finf : Rv, &m->n[f * nd], fl);
    const long o = m->owner[f], nbh = m->nbr[f];
    for (int v = 0; v < nv; v++) {   // Solver.cpp:41-67 (first-order states)
      L[v] = U[o * nv + v];
      Rv[v] = nbh >= 0 ? U[nbh * nv + v] : 0.0;
    }
    const double *ur = bt == BT_FREESTREAM ? finf : Rv;
    for (int v = 0; v < nv + 1; v++)
      URb[v] = (bt == BT_INTERNAL && v < nv) ? ur[v] : (bt == BT_FREESTREAM ? ur[v] : 0.0);
    jacobianOne(nd, bt, L, URb, &m->n[f * nd], gas, JL, JR);
    const double A = m->area[f];
    const double mL = dt[o] / m->vol[o] * A;
    for (int v = 0; v < nv; v++) {
      atomic_add(&s->b[o * nv + v], -(mL * fl[v]));
      atomic_add(&s->R[o * nv + v], fl[v] * A);
    }
    add(o, o, mL, JL);
    if (nbh >= 0) {
      const double mR = dt[nbh] / m->vol[nbh] * A;
      for (int v = 0; v < nv; v++) {
        atomic_add(&s->b[nbh * nv + v], dt[nbh] / m->vol[nbh] * A * fl[v]);
        atomic_add(&s->R[nbh * nv + v], -(fl[v] * A));
      }
      add(o, nbh, mL, JR);
      add(nbh, o, -mR, JL);
      add(nbh, nbh, -mR, JR);
    }
  }
}

}  // namespace

extern "C" {

or_implicit *or_implicit_create(const or_mesh *m, int restart, int maxiter, double rtol, int nthreads) {
  or_implicit *s = new or_implicit();
  const long ne = m->ncell;
  const int k = m->k, nv = m->nv, nb = nv * nv;
  s->nv = nv; s->k = k; s->ne = ne; s->restart = restart < 1 ? 1 : restart; s->maxiter = maxiter;
  s->rtol = rtol; s->nthreads = nthreads;
  s->rowptr.assign(ne + 1, 0);
  s->slot.assign(ne * k, -1);
  for (long e = 0; e < ne; e++) {
    s->rowptr[e] = (long)s->bcol.size();
    s->bcol.push_back(e);
    for (int mm = 0; mm < k; mm++) {
      const long j = m->cadj[e * k + mm];
      if (j >= 0) { s->slot[e * k + mm] = (long)s->bcol.size(); s->bcol.push_back(j); }
    }
  }
  s->rowptr[ne] = s->nnzb = (long)s->bcol.size();
  s->color.assign(ne, -1);
  int ncol = 0;
  for (long e = 0; e < ne; e++) {   // greedy colouring in cell order
    unsigned used = 0;
    for (int mm = 0; mm < k; mm++) {
      const long j = m->cadj[e * k + mm];
      if (j >= 0 && s->color[j] >= 0) used |= 1u << s->color[j];
    }
    int c = 0;
    while (used & (1u << c)) c++;
    s->color[e] = c;
    ncol = std::max(ncol, c + 1);
  }
  s->rows.assign(ncol, {});
  for (long e = 0; e < ne; e++) s->rows[s->color[e]].push_back(e);
  const long n = ne * nv;
  s->A.assign(s->nnzb * nb, 0.0);
  s->Dinv.assign(ne * nb, 0.0);
  s->b.assign(n, 0.0); s->R.assign(n, 0.0); s->x.assign(n, 0.0); s->dt.assign(ne, 0.0);
  s->V.assign((size_t)(s->restart + 1) * n, 0.0);
  s->Zv.assign(2 * n, 0.0); s->w.assign(n, 0.0); s->u.assign(n, 0.0);
  return s;
}

void or_implicit_destroy(or_implicit *s) { delete s; }

int or_implicit_ncolor(const or_implicit *s) { return (int)s->rows.size(); }

void or_implicit_freeze_limiter(or_implicit *s, const double *phi) {
  if (phi) s->phi_frozen.assign(phi, phi + s->ne * s->nv);
  else s->phi_frozen.clear();
}

int or_implicit_steps(or_implicit *s, const or_mesh *m, double *U, int nsteps, double *cfl, double cfl_min,
                      double cfl_max, const double *ex, int local, int repeat_nv, double stop_drop, double r0,
                      const double *finf, const or_gas *gas, double *lb_old, double *hist) {
  const int nv = m->nv;
  const long ne = m->ncell, n = ne * nv;
#ifdef _OPENMP
  const int prev = omp_get_max_threads();
  if (s->nthreads > 0) omp_set_num_threads(s->nthreads);
#endif
  int done = 0;
  for (int it = 0; it < nsteps; it++) {
    imp_assemble(m, s, U, *cfl, local, finf, gas, repeat_nv);
    imp_factor(m, s);
    double rr = 0;
    const long iters = imp_gmres(m, s, &rr);
    std::vector<double> lb(nv, 0.0), lr(nv, 0.0);
    for (long e = 0; e < ne; e++)
      for (int v = 0; v < nv; v++) {
        lb[v] = std::max(lb[v], std::fabs(s->b[e * nv + v]));
        lr[v] = std::max(lr[v], std::fabs(s->R[e * nv + v]));
      }
#pragma omp parallel for
    for (long e = 0; e < n; e++) U[e] += s->x[e];
    if (hist) {
      double *h = hist + (size_t)it * (2 * nv + 3);
      for (int v = 0; v < nv; v++) { h[v] = lb[v]; h[nv + v] = lr[v]; }
      h[2 * nv] = *cfl;
      h[2 * nv + 1] = (double)iters;
      h[2 * nv + 2] = rr;
    }
    done = it + 1;
    if (r0 <= 0) r0 = lr[0];
    if (cfl_max > cfl_min) {   // SolverImplicit.cpp:530-550
      const int ei = (lb_old[0] > lb[0]) ? 1 : 0;
      double c = *cfl * std::pow(lb_old[0] / lb[0], ex[ei]);
      c = std::max(cfl_min, c);
      *cfl = std::min(cfl_max, c);
    }
    std::copy(lb.begin(), lb.end(), lb_old);
    if (stop_drop > 0 && lr[0] <= stop_drop * r0) break;
  }
#ifdef _OPENMP
  omp_set_num_threads(prev);
#endif
  return done;
}

/* the system of the last step (BSR as or_implicit_assemble) and its solution */
void or_implicit_last(const or_implicit *s, double *bval, double *rhs, double *x) {
  if (bval) std::copy(s->A.begin(), s->A.end(), bval);
  if (rhs) std::copy(s->b.begin(), s->b.end(), rhs);
  if (x) std::copy(s->x.begin(), s->x.end(), x);
}

} // extern "C"
