// Golden-vector driver for the reference BES implementation.
//
// TEST INFRASTRUCTURE ONLY.  This program is linked against the reference's
// own Boost-free translation units (compiled from the reference tree by
// oracle/Makefile into oracle/_ref/) and dumps what those units compute, so
// that the CPU restatement in oracle/bes_oracle.cpp can be pinned against the
// reference itself.  It never ships and nothing on the GPU path uses it.
//
// The reference's Solver*.cpp and CoefficientMatrix.cpp cannot be compiled
// here (they include AMGCL, which needs Boost 1.66, absent from this image).
// The few lines of their face loops that the golden vectors need are
// therefore restated below on top of the reference objects:
//   * face reconstruction          Solver.cpp:69-121
//   * explicit stage face loop     SolverExplicit.cpp:275-378
//   * implicit face loop           SolverImplicit.cpp:358-507
//   * CSR pattern / block add      CoefficientMatrix.cpp:114-204
//   * AD Jacobian adapter          FluxJacobianWrapper.tpp:32-82
//
// Usage: bes_ref_driver <config.besi> <outdir>
//   always dumps geometry; dumps field quantities if <outdir>/state.bin exists;
//   runs explicit iterations if <outdir>/explicit_iters.txt exists;
//   evaluates flux / Jacobian known-answer inputs if <outdir>/kat_in.bin exists.

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "Config.hpp"
#include "Mesh.hpp"
#include "GasPerfect.hpp"
#include "SolutionSnapshot.hpp"
#include "GradientGreenGauss.hpp"
#include "LimiterVenkatakrishnan.hpp"
#include "TimeStepCalculatorLocal.hpp"
#include "TimeStepCalculatorGlobal.hpp"
#include "FluxRoe.hpp"
#include "FluxFreestream.hpp"
#include "FluxSlipWall.hpp"
#include "FluxWrapper.hpp"
#include "MatrixSparse.hpp"
#include <unsupported/Eigen/AutoDiff>

template <typename T>
static void dump(const std::string &path, const std::vector<T> &v) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) { perror(path.c_str()); exit(2); }
  if (!v.empty()) fwrite(v.data(), sizeof(T), v.size(), f);
  fclose(f);
}

static bool load(const std::string &path, std::vector<double> &v) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return false;
  fseek(f, 0, SEEK_END);
  long n = ftell(f) / (long)sizeof(double);
  fseek(f, 0, SEEK_SET);
  v.resize(n);
  size_t got = fread(v.data(), sizeof(double), n, f);
  fclose(f);
  return (long)got == n;
}

// AD Jacobian of a flux wrapper, as the reference adapter does it.
template <typename W>
static void adJacobian(Config *cfg, const double *left, const double *right,
                       const double *normal, double *jl, double *jr) {
  const double dummy[1] = {0};
  const double *rv = right ? right : dummy;
  Eigen::Matrix<double, W::InputsAtCompileTime, 1> x;
  Eigen::Matrix<double, W::ValuesAtCompileTime, 1> v;
  Eigen::Matrix<double, W::ValuesAtCompileTime, W::InputsAtCompileTime> jac;
  for (int i = 0; i < W::LeftInputAtCompileTime; i++) x[i] = left[i];
  for (int i = 0; i < W::RightInputAtCompileTime; i++) x[W::LeftInputAtCompileTime + i] = rv[i];
  Eigen::AutoDiffJacobian<W> fad;
  fad.initialize(cfg);
  fad.setNonDimNormal(normal);
  fad(x, &v, &jac);
  for (int o = 0; o < W::ValuesAtCompileTime; o++)
    for (int i = 0; i < W::LeftInputAtCompileTime; i++)
      jl[o * W::LeftInputAtCompileTime + i] = jac(o, i);
  if (jr)
    for (int o = 0; o < W::ValuesAtCompileTime; o++)
      for (int i = 0; i < W::RightInputAtCompileTime; i++)
        jr[o * W::RightInputAtCompileTime + i] = jac(o, i + W::LeftInputAtCompileTime);
}

struct Ctx {
  Config *cfg;
  Mesh *mesh;
  GasPerfect *gas;
  int nd, nv;
};

// Jacobian dispatch on (boundary type, dimension).
static void faceJacobian(Ctx &c, BoundaryType bt, const double *l, const double *r,
                         const double *n, double *jl, double *jr) {
  if (c.nd == 2) {
    if (bt == BoundaryType::INTERNAL)
      adJacobian<FluxWrapper_<FluxRoe_, 4, 4, 4> >(c.cfg, l, r, n, jl, jr);
    else if (bt == BoundaryType::FREESTREAM)
      adJacobian<FluxWrapper_<FluxFreestream_, 4, 5, 4> >(c.cfg, l, r, n, jl, NULL);
    else
      adJacobian<FluxWrapper_<FluxSlipWall_, 4, 1, 4> >(c.cfg, l, NULL, n, jl, NULL);
  } else {
    if (bt == BoundaryType::INTERNAL)
      adJacobian<FluxWrapper_<FluxRoe_, 5, 5, 5> >(c.cfg, l, r, n, jl, jr);
    else if (bt == BoundaryType::FREESTREAM)
      adJacobian<FluxWrapper_<FluxFreestream_, 5, 6, 5> >(c.cfg, l, r, n, jl, NULL);
    else
      adJacobian<FluxWrapper_<FluxSlipWall_, 5, 1, 5> >(c.cfg, l, NULL, n, jl, NULL);
  }
}

// Second-order face states from the reference objects (restates Solver.cpp:69-121).
static void faceStates(Ctx &c, SolutionSnapshot *s, Gradient *g, Limiter *lim,
                       unsigned long iFace, double *L, double *R) {
  Face *face = c.mesh->getFace(iFace);
  Element *e = face->getOwnerElement();
  for (int v = 0; v < c.nv; v++) {
    double phi = lim->getLimiterValue(e->getId(), v);
    L[v] = s->getConservedVariable(e->getId(), v);
    for (int d = 0; d < c.nd; d++)
      L[v] += phi * g->getGradient(e->getId(), v, d) * e->getVectorToFaceFromFaceId(iFace, d);
  }
  e = face->getNeighborElement();
  for (int v = 0; v < c.nv; v++) {
    if (e) {
      double phi = lim->getLimiterValue(e->getId(), v);
      R[v] = s->getConservedVariable(e->getId(), v);
      for (int d = 0; d < c.nd; d++)
        R[v] += phi * g->getGradient(e->getId(), v, d) * e->getVectorToFaceFromFaceId(iFace, d);
    } else {
      R[v] = 0;
    }
  }
}

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s <config.besi> <outdir>\n", argv[0]);
    return 2;
  }
  std::string out = argv[2];
  Ctx c;
  c.cfg = new Config();
  c.mesh = new Mesh();
  c.gas = new GasPerfect();
  c.cfg->initialize(argv[1]);
  c.mesh->initialize(c.cfg);
  c.cfg->initializePostMesh(c.mesh);
  c.gas->initialize(c.cfg);
  c.nd = c.mesh->getNumDim();
  c.nv = c.cfg->getNumConservedVariable();
  const int nd = c.nd, nv = c.nv;
  const unsigned long nf = c.mesh->getNumFace(), ne = c.mesh->getNumElement();
  const int k = c.mesh->getElement(0)->getNumFace();

  // ---------------- geometry ----------------
  {
    std::vector<long> owner(nf), nbr(nf), btype(nf);
    std::vector<double> S(nf * nd), n(nf * nd), area(nf), center(nf * nd);
    for (unsigned long f = 0; f < nf; f++) {
      Face *fc = c.mesh->getFace(f);
      owner[f] = fc->getOwnerElement()->getId();
      nbr[f] = fc->getNeighborElement() ? (long)fc->getNeighborElement()->getId() : -1;
      btype[f] = (long)fc->getBoundaryType();
      area[f] = fc->getArea();
      for (int d = 0; d < nd; d++) {
        S[f * nd + d] = fc->getDimNormal(d);
        n[f * nd + d] = fc->getNonDimNormal(d);
        center[f * nd + d] = fc->getCenter()->getSingleCoord(d);
      }
    }
    dump(out + "/face_owner.bin", owner);
    dump(out + "/face_nbr.bin", nbr);
    dump(out + "/face_btype.bin", btype);
    dump(out + "/face_S.bin", S);
    dump(out + "/face_n.bin", n);
    dump(out + "/face_area.bin", area);
    dump(out + "/face_center.bin", center);
    std::vector<double> vol(ne), cen(ne * nd), w(ne * k), r(ne * k * nd);
    std::vector<long> cf(ne * k), adj(ne * k), rsel(ne * k);
    for (unsigned long e = 0; e < ne; e++) {
      Element *el = c.mesh->getElement(e);
      vol[e] = el->getVolume();
      for (int d = 0; d < nd; d++) cen[e * nd + d] = el->getCentroid()->getSingleCoord(d);
      for (int m = 0; m < k; m++) {
        cf[e * k + m] = el->getSingleFace(m)->getId();
        adj[e * k + m] = el->getFaceAdjacentElement(m) ? (long)el->getFaceAdjacentElement(m)->getId() : -1;
        w[e * k + m] = el->getFaceInverseDistanceWeight(m);
        for (int d = 0; d < nd; d++) r[(e * k + m) * nd + d] = el->getVectorToFace(m)[d];
      }
    }
    // Which local vector does the face-id lookup used by the reconstruction
    // (Element::getVectorToFaceFromFaceId) return?  Identified by pointer.
    for (unsigned long e = 0; e < ne; e++) {
      Element *el = c.mesh->getElement(e);
      for (int m = 0; m < k; m++) {
        double *p = el->getVectorToFaceFromFaceId(el->getSingleFace(m)->getId());
        long sel = -1;
        for (int q = 0; q < k; q++) if (el->getVectorToFace(q) == p) sel = q;
        rsel[e * k + m] = sel;
      }
    }
    dump(out + "/cell_vol.bin", vol);
    dump(out + "/cell_centroid.bin", cen);
    dump(out + "/cell_faces.bin", cf);
    dump(out + "/cell_adj.bin", adj);
    dump(out + "/cell_w.bin", w);
    dump(out + "/cell_r.bin", r);
    dump(out + "/cell_rsel.bin", rsel);
    std::vector<long> meta = {nd, nv, (long)ne, (long)nf, k};
    dump(out + "/meta.bin", meta);
  }

  // ---------------- field quantities ----------------
  std::vector<double> st;
  if (load(out + "/state.bin", st) && st.size() == ne * nv) {
    SolutionSnapshot snap;
    snap.initialize(c.cfg, c.mesh, c.gas);
    for (unsigned long e = 0; e < ne; e++) snap.setConservedVariable(e, &st[e * nv]);
    GradientGreenGauss grad;
    grad.initialize(c.cfg, c.mesh, VariableType::CONSERVED);
    grad.computeGradient(&snap);
    LimiterVenkatakrishnan lim;
    lim.initialize(c.cfg, c.mesh, VariableType::CONSERVED);
    lim.computeLimiterTerm(&snap, &grad);
    TimeStepCalculatorLocal tsl;
    tsl.initialize(c.cfg, c.gas, c.mesh);
    tsl.computeAllTimeStep(&snap, c.cfg->getCourantNum());
    TimeStepCalculatorGlobal tsg;
    tsg.initialize(c.cfg, c.gas, c.mesh);
    tsg.computeAllTimeStep(&snap, c.cfg->getCourantNum());
    std::vector<double> G(ne * nv * nd), P(ne * nv), dtl(ne), dtg(1);
    for (unsigned long e = 0; e < ne; e++) {
      for (int v = 0; v < nv; v++) {
        P[e * nv + v] = lim.getLimiterValue(e, v);
        for (int d = 0; d < nd; d++) G[(e * nv + v) * nd + d] = grad.getGradient(e, v, d);
      }
      dtl[e] = tsl.getTimeStep(e);
    }
    dtg[0] = tsg.getTimeStep(0);
    dump(out + "/grad.bin", G);
    dump(out + "/lim.bin", P);
    dump(out + "/dt_local.bin", dtl);
    dump(out + "/dt_global.bin", dtg);

    FluxRoe roe; roe.initialize(c.cfg, c.gas);
    FluxFreestream fs; fs.initialize(c.cfg, c.gas);
    FluxSlipWall sw; sw.initialize(c.cfg, c.gas);
    double *finf = c.cfg->getFreestreamPrimitiveVariables(0);
    std::vector<double> R(ne * nv, 0.0), Fa(ne * nv, 0.0), F(nf * nv);
    double L[VAR_BUFFER], Rv[VAR_BUFFER], fl[VAR_BUFFER];
    // explicit face loop (SolverExplicit.cpp:285-344), serial, face order
    for (unsigned long f = 0; f < nf; f++) {
      Face *fc = c.mesh->getFace(f);
      faceStates(c, &snap, &grad, &lim, f, L, Rv);
      BoundaryType bt = fc->getBoundaryType();
      if (bt == BoundaryType::INTERNAL) roe.computeFlux(L, Rv, fc->getNonDimNormal(), fl);
      else if (bt == BoundaryType::FREESTREAM) fs.computeFlux(L, finf, fc->getNonDimNormal(), fl);
      else sw.computeFlux(L, NULL, fc->getNonDimNormal(), fl);
      unsigned long o = fc->getOwnerElement()->getId();
      for (int v = 0; v < nv; v++) {
        F[f * nv + v] = fl[v];
        R[o * nv + v] += fl[v] * fc->getArea();
        Fa[o * nv + v] += std::abs(fl[v] * fc->getArea());
      }
      if (fc->getNeighborElement()) {
        unsigned long nb = fc->getNeighborElement()->getId();
        for (int v = 0; v < nv; v++) {
          R[nb * nv + v] -= fl[v] * fc->getArea();
          Fa[nb * nv + v] += std::abs(fl[v] * fc->getArea());
        }
      }
    }
    dump(out + "/face_flux.bin", F);
    dump(out + "/resid.bin", R);
    dump(out + "/resid_scale.bin", Fa);

    // implicit assembly (SolverImplicit.cpp:347-507 with CoefficientMatrix.cpp:114-204)
    MatrixSparse mat;
    unsigned long nnz = 0;
    for (unsigned long e = 0; e < ne; e++) {
      nnz++;
      for (int m = 0; m < k; m++) if (c.mesh->getElement(e)->getFaceAdjacentElement(m)) nnz++;
    }
    nnz *= nv * nv;
    mat.initialize(ne * nv, nnz);
    for (unsigned long e = 0; e < ne; e++) {
      for (int a = 0; a < nv; a++) {
        for (int b = 0; b < nv; b++) mat.insertVal(e * nv + b, 0);
        for (int m = 0; m < k; m++) {
          Element *nb = c.mesh->getElement(e)->getFaceAdjacentElement(m);
          if (nb) for (int b = 0; b < nv; b++) mat.insertVal(nb->getId() * nv + b, 0);
        }
        mat.closeRow();
      }
    }
    mat.changeAllVal(0);
    for (unsigned long i = 0; i < ne * nv; i++) mat.addToVal(i, i, 1.0);
    auto addBlock = [&](unsigned long r0, unsigned long c0, double mul, const double *J) {
      for (int a = 0; a < nv; a++)
        for (int b = 0; b < nv; b++)
          mat.addToVal(r0 * nv + a, c0 * nv + b, mul * J[a * nv + b]);
    };
    std::vector<double> rhs(ne * nv, 0.0), JLall(nf * nv * nv), JRall(nf * nv * nv, 0.0);
    double JL[VAR_BUFFER * VAR_BUFFER], JR[VAR_BUFFER * VAR_BUFFER];
    for (unsigned long f = 0; f < nf; f++) {
      Face *fc = c.mesh->getFace(f);
      BoundaryType bt = fc->getBoundaryType();
      faceStates(c, &snap, &grad, &lim, f, L, Rv);
      if (bt == BoundaryType::INTERNAL) roe.computeFlux(L, Rv, fc->getNonDimNormal(), fl);
      else if (bt == BoundaryType::FREESTREAM) fs.computeFlux(L, finf, fc->getNonDimNormal(), fl);
      else sw.computeFlux(L, NULL, fc->getNonDimNormal(), fl);
      // first-order states (Solver.cpp:41-67)
      Element *oe = fc->getOwnerElement(), *ne_ = fc->getNeighborElement();
      for (int v = 0; v < nv; v++) {
        L[v] = snap.getConservedVariable(oe->getId(), v);
        Rv[v] = ne_ ? snap.getConservedVariable(ne_->getId(), v) : 0.0;
      }
      const double *rightIn = (bt == BoundaryType::INTERNAL) ? Rv : (bt == BoundaryType::FREESTREAM ? finf : NULL);
      faceJacobian(c, bt, L, rightIn, fc->getNonDimNormal(), JL, bt == BoundaryType::INTERNAL ? JR : NULL);
      for (int q = 0; q < nv * nv; q++) {
        JLall[f * nv * nv + q] = JL[q];
        if (bt == BoundaryType::INTERNAL) JRall[f * nv * nv + q] = JR[q];
      }
      unsigned long o = oe->getId();
      double mL = tsl.getTimeStep(o) / oe->getVolume() * fc->getArea();
      for (int v = 0; v < nv; v++) {
        rhs[o * nv + v] -= mL * fl[v];
        addBlock(o, o, mL, JL);
      }
      if (ne_) {
        unsigned long nb = ne_->getId();
        double mR = tsl.getTimeStep(nb) / ne_->getVolume() * fc->getArea();
        for (int v = 0; v < nv; v++) {
          rhs[nb * nv + v] += tsl.getTimeStep(nb) / ne_->getVolume() * fc->getArea() * fl[v];
          addBlock(o, nb, mL, JR);
          addBlock(nb, o, -mR, JL);
          addBlock(nb, nb, -mR, JR);
        }
      }
    }
    std::vector<long> rp(mat.getRowPtr(), mat.getRowPtr() + ne * nv + 1);
    std::vector<long> ci(mat.getColPtr(), mat.getColPtr() + mat.getNumNonZero());
    std::vector<double> va(mat.getValPtr(), mat.getValPtr() + mat.getNumNonZero());
    dump(out + "/csr_rowptr.bin", rp);
    dump(out + "/csr_col.bin", ci);
    dump(out + "/csr_val.bin", va);
    dump(out + "/impl_rhs.bin", rhs);
    dump(out + "/face_JL.bin", JLall);
    dump(out + "/face_JR.bin", JRall);

    // explicit multi-stage iterations (SolverExplicit.cpp:259-393), serial
    std::ifstream itf(out + "/explicit_iters.txt");
    long niter = 0;
    if (itf >> niter && niter > 0) {
      SolutionSnapshot s0, s1;
      s0.initialize(c.cfg, c.mesh, c.gas);
      s1.initialize(c.cfg, c.mesh, c.gas);
      for (unsigned long e = 0; e < ne; e++) {
        s0.setConservedVariable(e, &st[e * nv]);
        s1.setConservedVariable(e, &st[e * nv]);
      }
      SolutionSnapshot *cur0 = &s0, *cur1 = &s1;
      std::vector<double> linf;
      std::vector<double> sub(ne * nv);
      for (long it = 0; it < niter; it++) {
        tsl.computeAllTimeStep(cur0, c.cfg->getCourantNum());
        for (int iSub = 0; iSub < c.cfg->getNumSubTimeStep(); iSub++) {
          std::fill(sub.begin(), sub.end(), 0.0);
          SolutionSnapshot *ss = iSub == 0 ? cur0 : cur1;
          grad.computeGradient(ss);
          lim.computeLimiterTerm(ss, &grad);
          for (unsigned long f = 0; f < nf; f++) {
            Face *fc = c.mesh->getFace(f);
            faceStates(c, ss, &grad, &lim, f, L, Rv);
            BoundaryType bt = fc->getBoundaryType();
            if (bt == BoundaryType::INTERNAL) roe.computeFlux(L, Rv, fc->getNonDimNormal(), fl);
            else if (bt == BoundaryType::FREESTREAM) fs.computeFlux(L, finf, fc->getNonDimNormal(), fl);
            else sw.computeFlux(L, NULL, fc->getNonDimNormal(), fl);
            unsigned long o = fc->getOwnerElement()->getId();
            for (int v = 0; v < nv; v++) sub[o * nv + v] += fl[v] * fc->getArea();
            if (fc->getNeighborElement()) {
              unsigned long nb = fc->getNeighborElement()->getId();
              for (int v = 0; v < nv; v++) sub[nb * nv + v] -= fl[v] * fc->getArea();
            }
          }
          bool last = iSub == c.cfg->getNumSubTimeStep() - 1;
          std::vector<double> mx(nv, 0.0);
          for (unsigned long e = 0; e < ne; e++) {
            for (int v = 0; v < nv; v++) {
              if (last) mx[v] = std::abs(sub[e * nv + v]) > mx[v] ? std::abs(sub[e * nv + v]) : mx[v];
              double nu = cur0->getConservedVariable(e, v) -
                (c.cfg->getTimeStepConstant(iSub) * tsl.getTimeStep(e) / c.mesh->getElement(e)->getVolume()) *
                sub[e * nv + v];
              cur1->setConservedVariable(e, v, nu);
            }
          }
          if (last) linf.insert(linf.end(), mx.begin(), mx.end());
        }
        std::swap(cur0, cur1);
      }
      std::vector<double> fin(ne * nv);
      for (unsigned long e = 0; e < ne; e++)
        for (int v = 0; v < nv; v++) fin[e * nv + v] = cur0->getConservedVariable(e, v);
      dump(out + "/explicit_state.bin", fin);
      dump(out + "/explicit_linf.bin", linf);
    }
  }

  // ---------------- flux / Jacobian known answers ----------------
  std::vector<double> kin;
  if (load(out + "/kat_in.bin", kin)) {
    // record layout: [kind, UL(nv), UR(nv+1), n(nd)], kind 1=Roe 2=slip 3=freestream
    const int rec = 1 + nv + (nv + 1) + nd;
    const long ns = (long)kin.size() / rec;
    FluxRoe roe; roe.initialize(c.cfg, c.gas);
    FluxFreestream fs; fs.initialize(c.cfg, c.gas);
    FluxSlipWall sw; sw.initialize(c.cfg, c.gas);
    std::vector<double> F(ns * nv), JL(ns * nv * nv), JR(ns * nv * nv, 0.0);
    for (long s = 0; s < ns; s++) {
      const double *x = &kin[s * rec];
      int kind = (int)x[0];
      const double *ul = x + 1, *ur = x + 1 + nv, *n = x + 1 + nv + nv + 1;
      BoundaryType bt = kind == 1 ? BoundaryType::INTERNAL :
                        (kind == 2 ? BoundaryType::SLIP_WALL : BoundaryType::FREESTREAM);
      if (kind == 1) roe.computeFlux(ul, ur, n, &F[s * nv]);
      else if (kind == 2) sw.computeFlux(ul, NULL, n, &F[s * nv]);
      else fs.computeFlux(ul, ur, n, &F[s * nv]);
      faceJacobian(c, bt, ul, kind == 2 ? NULL : ur, n, &JL[s * nv * nv],
                   kind == 1 ? &JR[s * nv * nv] : NULL);
    }
    dump(out + "/kat_flux.bin", F);
    dump(out + "/kat_JL.bin", JL);
    dump(out + "/kat_JR.bin", JR);
  }
  return 0;
}
