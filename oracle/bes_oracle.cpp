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
