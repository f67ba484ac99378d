// Explicit residual path of the BES solver core on MI355X (gfx950).
//
// Per pseudo-time step (SolverExplicit.cpp:259-393):
//   k_timestep   local dt per cell + global minimum (TimeStepCalculator*.cpp)
//   per stage:
//     k_gradlim  Green-Gauss gradient fused with the Venkatakrishnan limiter,
//                writes the limited gradient phi*grad (GradientGreenGauss.cpp:21-135,
//                LimiterVenkatakrishnan.cpp:21-128)
//     k_residual second-order face fluxes gathered per cell, summed in the
//                reference face order, fused with the stage update
//                U = U0 - alpha*dt/V*R and the Linf monitor
//                (SolverExplicit.cpp:285-378, Solver.cpp:69-121)
// One thread per cell; every cell computes the fluxes of its own faces, so no
// atomics touch the residual and the result is deterministic.  Blocks are
// remapped so that each XCD walks a contiguous range of Morton-ordered cells.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bes_hip.h"
#include "context.hpp"
#include "flux.hpp"
#include "kernels_common.hpp"

namespace bes {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
int hip_fail(hipError_t e, const char *what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return BES_ERR_HIP;
}

static hipEvent_t take_event(bes_ctx *c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

ProfScope::ProfScope(bes_ctx *c_, int kind_, double bytes_) : c(c_), kind(kind_), bytes(bytes_) {
  if (!c->prof) return;
  a = take_event(c);
  b = take_event(c);
  if (a) hipEventRecord(a, c->stream);
}
ProfScope::~ProfScope() {
  if (!c->prof || !a || !b) return;
  hipEventRecord(b, c->stream);
  c->ev_pending.push_back({kind, bytes, a, b});
}

int prof_collect(bes_ctx *c) {
  if (c->ev_pending.empty()) return BES_OK;
  BES_HIP(hipStreamSynchronize(c->stream));
  for (auto &p : c->ev_pending) {
    float ms = 0;
    BES_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    c->prof_ms[p.kind] += ms;
    c->prof_n[p.kind] += 1;
    c->prof_bytes[p.kind] += p.bytes;
    c->ev_pool.push_back(p.a);
    c->ev_pool.push_back(p.b);
  }
  c->ev_pending.clear();
  return BES_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

template <int ND>
__global__ __launch_bounds__(256) void k_timestep(const double *__restrict__ U, DevMesh dm,
                                                  GasConst gc, double cfl,
                                                  double *__restrict__ dt,
                                                  unsigned long long *__restrict__ gmin) {
  constexpr int NV = ND + 2;
  const int N = dm.N;
  const int i = cell_index();
  double t = __builtin_inf();
  if (i < N) {
    double u[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) u[v] = U[(size_t)v * N + i];
    Gas<ND, double> g(gc);
    double c = g.sound_speed(u);
    double sr = 0.0;
#pragma unroll
    for (int d = 0; d < ND; d++) sr += (c + ::fabs(g.velocity(u, d))) * dm.sproj[(size_t)d * N + i];
    t = cfl * dm.vol[i] / sr;
    dt[i] = t;
  }
  if (gmin) {
    double m = block_min(t);
    if (threadIdx.x == 0) atomicMin(gmin, (unsigned long long)__double_as_longlong(m));
  }
}

__global__ void k_decode_min(const unsigned long long *bits, double *out) {
  out[0] = __longlong_as_double((long long)bits[0]);
}

// reference LimiterVenkatakrishnan.cpp:101-128
__device__ __forceinline__ double face_limiter(double mn, double mx, double d2, double eps2) {
  if (d2 > 0)
    return (1 / d2) * ((mx * mx + eps2) * d2 + 2.0 * d2 * d2 * mx) /
           (mx * mx + 2.0 * d2 * d2 + mx * d2 + eps2);
  else if (d2 < 0)
    return (1 / d2) * ((mn * mn + eps2) * d2 + 2.0 * d2 * d2 * mn) /
           (mn * mn + 2.0 * d2 * d2 + mn * d2 + eps2);
  return 1.0;
}

template <int ND>
__global__ __launch_bounds__(256) void k_gradlim(const double *__restrict__ U, DevMesh dm,
                                                 double *__restrict__ LG,
                                                 double *__restrict__ Gout,
                                                 double *__restrict__ Pout) {
  constexpr int NV = ND + 2, K = ND + 1;
  const int N = dm.N;
  const int i = cell_index();
  if (i >= N) return;
  double u[NV], mn[NV], mx[NV], g[NV][ND];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    u[v] = U[(size_t)v * N + i];
    mn[v] = u[v];
    mx[v] = u[v];
#pragma unroll
    for (int d = 0; d < ND; d++) g[v][d] = 0.0;
  }
#pragma unroll
  for (int m = 0; m < K; m++) {
    const int j = dm.nbr[(size_t)m * N + i];
    double S[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) S[d] = dm.Sg[((size_t)m * ND + d) * N + i];
    if (j >= 0) {
      const double w = dm.w[(size_t)m * N + i];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        const double uj = U[(size_t)v * N + j];
        const double fv = w * u[v] + (1 - w) * uj;
#pragma unroll
        for (int d = 0; d < ND; d++) g[v][d] += S[d] * fv;
        mn[v] = (uj < mn[v]) ? uj : mn[v];
        mx[v] = (uj > mx[v]) ? uj : mx[v];
      }
    } else {
#pragma unroll
      for (int v = 0; v < NV; v++)
#pragma unroll
        for (int d = 0; d < ND; d++) g[v][d] += S[d] * (u[v]);
      if (j == -1) {  // slip wall: tangential momentum on the face
        double nn[ND], t[ND], vn = 0;
#pragma unroll
        for (int d = 0; d < ND; d++) nn[d] = dm.nf[((size_t)m * ND + d) * N + i];
#pragma unroll
        for (int d = 0; d < ND; d++) vn += u[1 + d] * nn[d];
#pragma unroll
        for (int d = 0; d < ND; d++) t[d] = u[1 + d] - vn * nn[d];
#pragma unroll
        for (int a = 0; a < ND; a++)
#pragma unroll
          for (int d = 0; d < ND; d++) g[1 + a][d] += (t[a] * S[d] - u[1 + a] * S[d]);
      }
    }
  }
  const double V = dm.vol[i];
#pragma unroll
  for (int v = 0; v < NV; v++)
#pragma unroll
    for (int d = 0; d < ND; d++) g[v][d] /= V;
  const double eps2 = dm.eps2[i];
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double dmin = mn[v] - u[v], dmax = mx[v] - u[v];
    double best = 1.0;
#pragma unroll
    for (int m = 0; m < K; m++) {
      if (dm.nbr[(size_t)m * N + i] < 0) continue;
      double d2 = 0;
#pragma unroll
      for (int d = 0; d < ND; d++) d2 += g[v][d] * dm.rl[((size_t)m * ND + d) * N + i];
      d2 /= 2.0;
      const double f = face_limiter(dmin, dmax, d2, eps2);
      best = (f < best) ? f : best;
    }
#pragma unroll
    for (int d = 0; d < ND; d++) {
      LG[((size_t)v * ND + d) * N + i] = best * g[v][d];
      if (Gout) Gout[((size_t)v * ND + d) * N + i] = g[v][d];
    }
    if (Pout) Pout[(size_t)v * N + i] = best;
  }
}

struct Finf {
  double p[6];
};

template <int ND>
__global__ __launch_bounds__(256) void k_residual(
    // U0 and Ud may be the same array (the last stage updates in place: each
    // thread reads only its own U0[i] before writing Ud[i]), so no restrict
    const double *__restrict__ Us, const double *__restrict__ LG, const double *U0,
    double *Ud, const double *__restrict__ dt, const double *__restrict__ dtg,
    double alpha, DevMesh dm, GasConst gc, Finf fi, double *__restrict__ Rout,
    unsigned long long *__restrict__ linf_bits, unsigned long long *__restrict__ err) {
  constexpr int NV = ND + 2, K = ND + 1;
  const int N = dm.N;
  const int i = cell_index();
  double R[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) R[v] = 0.0;
  if (i < N) {
    double u[NV], lg[NV][ND];
#pragma unroll
    for (int v = 0; v < NV; v++) {
      u[v] = Us[(size_t)v * N + i];
#pragma unroll
      for (int d = 0; d < ND; d++) lg[v][d] = LG[((size_t)v * ND + d) * N + i];
    }
    const uint32_t info = dm.info[i];
    for (int q = 0; q < K; q++) {
      const int m = (info >> (2 * q)) & 3;
      const bool own = (info >> (8 + m)) & 1;
      const int j = dm.nbr[(size_t)m * N + i];
      double nn[ND], rs[ND];
#pragma unroll
      for (int d = 0; d < ND; d++) {
        nn[d] = dm.nf[((size_t)m * ND + d) * N + i];
        rs[d] = dm.rs[((size_t)m * ND + d) * N + i];
      }
      const double A = dm.af[(size_t)m * N + i];
      double ui[NV], F[NV];
#pragma unroll
      for (int v = 0; v < NV; v++) {
        ui[v] = u[v];
#pragma unroll
        for (int d = 0; d < ND; d++) ui[v] += lg[v][d] * rs[d];
      }
      if (j >= 0) {
        double rn[ND], uj[NV];
#pragma unroll
        for (int d = 0; d < ND; d++) rn[d] = dm.rn[((size_t)m * ND + d) * N + i];
#pragma unroll
        for (int v = 0; v < NV; v++) {
          uj[v] = Us[(size_t)v * N + j];
#pragma unroll
          for (int d = 0; d < ND; d++) uj[v] += LG[((size_t)v * ND + d) * N + j] * rn[d];
        }
        if (own) roe_flux<ND, double>(gc, ui, uj, nn, F);
        else roe_flux<ND, double>(gc, uj, ui, nn, F);
      } else if (j == -1) {
        slip_flux<ND, double>(gc, ui, nn, F);
      } else {
        freestream_flux<ND, double>(gc, ui, fi.p, nn, F);
      }
      bool bad = false;
#pragma unroll
      for (int v = 0; v < NV; v++) bad |= (F[v] != F[v]);
      if (bad) atomicMin(err, (unsigned long long)i * K + m);
#pragma unroll
      for (int v = 0; v < NV; v++) {
        if (own) R[v] += F[v] * A;
        else R[v] -= F[v] * A;
      }
    }
    if (Ud) {
      const double c = alpha * (dtg ? dtg[0] : dt[i]) / dm.vol[i];
#pragma unroll
      for (int v = 0; v < NV; v++) Ud[(size_t)v * N + i] = U0[(size_t)v * N + i] - c * R[v];
    }
    if (Rout) {
#pragma unroll
      for (int v = 0; v < NV; v++) Rout[(size_t)v * N + i] = R[v];
    }
  }
  if (linf_bits) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      double a = block_max(::fabs(R[v]));
      if (threadIdx.x == 0) atomicMax(&linf_bits[v], (unsigned long long)__double_as_longlong(a));
    }
  }
}

__global__ void k_fill_u64(unsigned long long *p, int n, unsigned long long v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_bits_to_double(const unsigned long long *bits, double *out, int n) {
  int i = threadIdx.x;
  if (i < n) out[i] = __longlong_as_double((long long)bits[i]);
}

// per-face flux / Jacobian batches (host-facing known-answer interface)
template <int ND>
__global__ void k_flux_batch(int kind, int n, const double *__restrict__ UL,
                             const double *__restrict__ UR, const double *__restrict__ nrm,
                             GasConst gc, double *__restrict__ F) {
  constexpr int NV = ND + 2;
  int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  double ul[NV], ur[NV + 1], nn[ND], out[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) ul[v] = UL[(size_t)f * NV + v];
#pragma unroll
  for (int v = 0; v < NV + 1; v++) ur[v] = UR[(size_t)f * (NV + 1) + v];
#pragma unroll
  for (int d = 0; d < ND; d++) nn[d] = nrm[(size_t)f * ND + d];
  if (kind == kInternal) roe_flux<ND, double>(gc, ul, ur, nn, out);
  else if (kind == kSlipWall) slip_flux<ND, double>(gc, ul, nn, out);
  else freestream_flux<ND, double>(gc, ul, ur, nn, out);
#pragma unroll
  for (int v = 0; v < NV; v++) F[(size_t)f * NV + v] = out[v];
}

// one thread per (face, input direction); 2*NV directions for Roe, NV otherwise
template <int ND>
__global__ void k_jacobian_batch(int kind, int n, const double *__restrict__ UL,
                                 const double *__restrict__ UR, const double *__restrict__ nrm,
                                 GasConst gc, double *__restrict__ JL, double *__restrict__ JR) {
  constexpr int NV = ND + 2;
  const int ndir = 2 * NV;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int f = t / ndir, dir = t % ndir;
  if (f >= n) return;
  if (kind != kInternal && dir >= NV) return;
  Dual ul[NV], ur[NV + 1], out[NV];
  double nn[ND];
#pragma unroll
  for (int v = 0; v < NV; v++) ul[v] = Dual(UL[(size_t)f * NV + v], dir == v ? 1.0 : 0.0);
#pragma unroll
  for (int v = 0; v < NV + 1; v++) ur[v] = Dual(UR[(size_t)f * (NV + 1) + v], dir == NV + v ? 1.0 : 0.0);
#pragma unroll
  for (int d = 0; d < ND; d++) nn[d] = nrm[(size_t)f * ND + d];
  if (kind == kInternal) roe_flux<ND, Dual>(gc, ul, ur, nn, out);
  else if (kind == kSlipWall) slip_flux<ND, Dual>(gc, ul, nn, out);
  else freestream_flux<ND, Dual>(gc, ul, ur, nn, out);
  double *J = dir < NV ? JL : JR;
  int col = dir < NV ? dir : dir - NV;
  if (!J) return;
#pragma unroll
  for (int o = 0; o < NV; o++) J[((size_t)f * NV + o) * NV + col] = out[o].d;
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

template <typename T>
static int dalloc(T **p, size_t n) {
  BES_HIP(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
  return BES_OK;
}

template <typename T>
static int upload(T **p, const std::vector<T> &h) {
  int s = dalloc(p, h.size());
  if (s) return s;
  if (!h.empty()) BES_HIP(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return BES_OK;
}

static GasConst make_gas(const bes_gas *gas) {
  GasConst g;
  g.gamma = gas ? gas->gamma : 1.4;
  g.R = gas ? gas->R : 287.0;
  g.harten = gas ? gas->harten : 0.05;
  g.cv = g.R / (g.gamma - 1.0);
  g.cp = g.cv + g.R;
  return g;
}

// Device layout from the reference geometry (cells in Morton order).
static int build_device_mesh(bes_ctx *c) {
  const Geometry &g = c->geo;
  const int nd = g.nd, k = g.k, N = (int)g.ncell;
  c->perm = locality_order(g);
  c->iperm.assign(N, 0);
  for (int i = 0; i < N; i++) c->iperm[c->perm[i]] = i;
  std::vector<double> vol(N), eps2(N), sproj((size_t)nd * N), Sg((size_t)k * nd * N), w((size_t)k * N),
      rl((size_t)k * nd * N), nf((size_t)k * nd * N), af((size_t)k * N), rs((size_t)k * nd * N),
      rn((size_t)k * nd * N);
  std::vector<int> nbr((size_t)k * N);
  std::vector<uint32_t> info(N);
  for (int i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    vol[i] = g.vol[e];
    // LimiterVenkatakrishnan.cpp:42-46, epsilon squared where it is used
    double L = std::pow(g.vol[e], 1.0 / nd);
    double eps = std::pow(c->venkat_k * L, (3.0 / 2.0));
    eps2[i] = eps * eps;
    for (int d = 0; d < nd; d++) sproj[(size_t)d * N + i] = g.sproj[e * nd + d];
    int order[4];
    for (int m = 0; m < k; m++) order[m] = m;
    std::sort(order, order + k, [&](int a, int b) { return g.cface[e * k + a] < g.cface[e * k + b]; });
    uint32_t bits = 0;
    for (int q = 0; q < k; q++) bits |= (uint32_t)order[q] << (2 * q);
    for (int m = 0; m < k; m++) {
      const int64_t s = e * k + m, f = g.cface[s];
      const bool own = g.owner[f] == e;
      if (own) bits |= 1u << (8 + m);
      const int64_t jadj = g.cadj[s];
      int code;
      if (jadj >= 0) code = (int)c->iperm[jadj];
      else code = (g.btype[f] == kSlipWall) ? -1 : -2;
      nbr[(size_t)m * N + i] = code;
      w[(size_t)m * N + i] = g.w[s];
      af[(size_t)m * N + i] = g.area[f];
      const int64_t sel = e * k + g.rsel[s];
      // neighbour's own reconstruction vector for this face
      int64_t nsel = -1;
      if (jadj >= 0)
        for (int mm = 0; mm < k; mm++)
          if (g.cface[jadj * k + mm] == f) nsel = jadj * k + g.rsel[jadj * k + mm];
      for (int d = 0; d < nd; d++) {
        const size_t q = ((size_t)m * nd + d) * N + i;
        Sg[q] = own ? g.S[f * nd + d] : -g.S[f * nd + d];
        rl[q] = g.r[s * nd + d];
        nf[q] = g.n[f * nd + d];
        rs[q] = g.r[sel * nd + d];
        rn[q] = nsel >= 0 ? g.r[nsel * nd + d] : 0.0;
      }
    }
    info[i] = bits;
  }
  DevMesh &dm = c->dm;
  dm.nd = nd; dm.k = k; dm.N = N;
  int s;
  if ((s = upload(&dm.vol, vol)) || (s = upload(&dm.eps2, eps2)) || (s = upload(&dm.sproj, sproj)) ||
      (s = upload(&dm.nbr, nbr)) || (s = upload(&dm.Sg, Sg)) || (s = upload(&dm.w, w)) ||
      (s = upload(&dm.rl, rl)) || (s = upload(&dm.nf, nf)) || (s = upload(&dm.af, af)) ||
      (s = upload(&dm.rs, rs)) || (s = upload(&dm.rn, rn)) || (s = upload(&dm.info, info)))
    return s;
  return BES_OK;
}

static int create_from_input(const MeshInput &in, int geom_mode, const bes_gas *gas, bes_ctx **out) {
  if (!out) { set_error("null output pointer"); return BES_ERR_ARG; }
  *out = nullptr;
  if (geom_mode < 0 || geom_mode > 7) { set_error("geom_mode must be in [0,7]"); return BES_ERR_ARG; }
  bes_ctx *c = new bes_ctx();
  try {
    c->geo = build_geometry(in, geom_mode);
  } catch (const std::exception &ex) {
    set_error(ex.what());
    delete c;
    return BES_ERR_MESH;
  }
  if (c->geo.ncell > 0x7fffffff / 4) { set_error("mesh too large"); delete c; return BES_ERR_ARG; }
  c->nd = c->geo.nd; c->nv = c->nd + 2; c->k = c->geo.k; c->N = (int)c->geo.ncell;
  c->gc = make_gas(gas);
  c->venkat_k = gas ? gas->venkat_k : 5.0;
  int s;
  hipError_t he = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (he != hipSuccess) { s = hip_fail(he, "hipStreamCreate"); delete c; return s; }
  const size_t N = c->N, nv = c->nv, nd = c->nd;
  if ((s = build_device_mesh(c)) || (s = dalloc(&c->U0, nv * N)) || (s = dalloc(&c->Ua, nv * N)) ||
      (s = dalloc(&c->Ub, nv * N)) || (s = dalloc(&c->LG, nv * nd * N)) || (s = dalloc(&c->dt, N)) ||
      (s = dalloc(&c->dtg, 1)) || (s = dalloc(&c->red, 64)) || (s = dalloc(&c->err, 1))) {
    bes_ctx_destroy(c);
    return s;
  }
  // default freestream: Mach 0 (Config.cpp:481-487 form)
  double fi[6] = {1.0, 0, 0, 0, 0, 0};
  fi[nd + 1] = 1.0 / c->gc.gamma;
  fi[nd + 2] = 1.0 / (c->gc.gamma * c->gc.R);
  std::memcpy(c->finf, fi, sizeof(fi));
  *out = c;
  return BES_OK;
}

static int reset_err(bes_ctx *c) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->err, 1, ~0ull);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

static int check_err(bes_ctx *c) {
  unsigned long long e = 0;
  BES_HIP(hipMemcpyAsync(&e, c->err, sizeof(e), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  if (e != ~0ull) {
    const int64_t cell = c->perm[e / c->k];
    const int m = (int)(e % c->k);
    c->last_err_face = (long)c->geo.cface[cell * c->k + m];
    set_error("non-finite flux at face " + std::to_string(c->last_err_face));
    return BES_ERR_NAN;
  }
  return BES_OK;
}

// launch helpers shared with the implicit path
int launch_timestep(bes_ctx *c, const double *U, double cfl, int local) {
  const int bs = 256, nb = grid_for(c->N, bs);
  unsigned long long *gm = nullptr;
  if (!local) {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->red, 1, 0x7ff0000000000000ull);
    gm = c->red;
  }
  {
    ProfScope ps(c, kProfTimestep, (double)c->N * 8 * (c->nv + c->nd + 2));
    if (c->nd == 2) hipLaunchKernelGGL(k_timestep<2>, dim3(nb), dim3(bs), 0, c->stream, U, c->dm, c->gc, cfl, c->dt, gm);
    else hipLaunchKernelGGL(k_timestep<3>, dim3(nb), dim3(bs), 0, c->stream, U, c->dm, c->gc, cfl, c->dt, gm);
  }
  if (!local) hipLaunchKernelGGL(k_decode_min, dim3(1), dim3(1), 0, c->stream, c->red, c->dtg);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

int launch_gradlim(bes_ctx *c, const double *U, double *Gout, double *Pout) {
  const int bs = 256, nb = grid_for(c->N, bs);
  ProfScope ps(c, kProfGradlim,
               (double)c->N * (8 * c->nv + c->k * (4 + 16 * c->nd + 8) + 16 + 8 * c->nv * c->nd));
  if (c->nd == 2) hipLaunchKernelGGL(k_gradlim<2>, dim3(nb), dim3(bs), 0, c->stream, U, c->dm, c->LG, Gout, Pout);
  else hipLaunchKernelGGL(k_gradlim<3>, dim3(nb), dim3(bs), 0, c->stream, U, c->dm, c->LG, Gout, Pout);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

int launch_residual(bes_ctx *c, const double *Us, const double *U0, double *Ud, double alpha,
                    int local, double *Rout, unsigned long long *linf_bits) {
  const int bs = 256, nb = grid_for(c->N, bs);
  Finf fi;
  std::memcpy(fi.p, c->finf, sizeof(fi.p));
  const double *dtg = local ? nullptr : c->dtg;
  ProfScope ps(c, kProfResidual,
               (double)c->N * (8 * c->nv + 8 * c->nv * c->nd + c->k * (12 + 24 * c->nd) + 20 +
                               (U0 ? 16 * c->nv : 0) + (Rout ? 8 * c->nv : 0)));
  if (c->nd == 2)
    hipLaunchKernelGGL(k_residual<2>, dim3(nb), dim3(bs), 0, c->stream, Us, c->LG, U0, Ud, c->dt, dtg,
                       alpha, c->dm, c->gc, fi, Rout, linf_bits, c->err);
  else
    hipLaunchKernelGGL(k_residual<3>, dim3(nb), dim3(bs), 0, c->stream, Us, c->LG, U0, Ud, c->dt, dtg,
                       alpha, c->dm, c->gc, fi, Rout, linf_bits, c->err);
  BES_HIP(hipGetLastError());
  return BES_OK;
}

// device SoA (Morton order) <-> host cell-major (reference order)
int state_to_host(bes_ctx *c, const double *dsoa, int ncomp, double *host) {
  const size_t N = c->N;
  std::vector<double> tmp(ncomp * N);
  BES_HIP(hipMemcpyAsync(tmp.data(), dsoa, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    for (int q = 0; q < ncomp; q++) host[e * ncomp + q] = tmp[(size_t)q * N + i];
  }
  return BES_OK;
}

int state_to_device(bes_ctx *c, const double *host, int ncomp, double *dsoa) {
  const size_t N = c->N;
  std::vector<double> tmp(ncomp * N);
  for (size_t i = 0; i < N; i++) {
    const int64_t e = c->perm[i];
    for (int q = 0; q < ncomp; q++) tmp[(size_t)q * N + i] = host[e * ncomp + q];
  }
  BES_HIP(hipMemcpyAsync(dsoa, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

}  // namespace bes

using namespace bes;

// ---------------------------------------------------------------------------
// C ABI (declared in include/bes_hip.h)
// ---------------------------------------------------------------------------
extern "C" {

const char *bes_last_error_message(void) { return g_err.c_str(); }

const char *bes_status_string(int s) {
  switch (s) {
    case BES_OK: return "ok";
    case BES_ERR_ARG: return "invalid argument";
    case BES_ERR_HIP: return "HIP runtime error";
    case BES_ERR_NAN: return "non-finite flux";
    case BES_ERR_MESH: return "mesh error";
    case BES_ERR_SOLVE: return "linear solver did not converge";
    default: return "unknown status";
  }
}

int bes_ctx_create(const bes_mesh_desc *mesh, int geom_mode, const bes_gas *gas, bes_ctx **out) {
  if (!mesh || !mesh->coords || !mesh->cells || (mesh->nbface > 0 && (!mesh->bfaces || !mesh->btype)) ||
      (mesh->ndim != 2 && mesh->ndim != 3) || mesh->ncell <= 0 || mesh->nnode <= 0) {
    set_error("invalid mesh descriptor");
    return BES_ERR_ARG;
  }
  MeshInput in;
  in.ndim = mesh->ndim;
  in.coords.assign(mesh->coords, mesh->coords + mesh->nnode * mesh->ndim);
  in.cells.assign(mesh->cells, mesh->cells + mesh->ncell * (mesh->ndim + 1));
  if (mesh->nbface > 0) {
    in.bfaces.assign(mesh->bfaces, mesh->bfaces + mesh->nbface * mesh->ndim);
    in.btype.assign(mesh->btype, mesh->btype + mesh->nbface);
  }
  return create_from_input(in, geom_mode, gas, out);
}

int bes_ctx_create_msh2(const char *path, int geom_mode, const bes_gas *gas, bes_ctx **out) {
  if (!path) { set_error("null path"); return BES_ERR_ARG; }
  MeshInput in;
  try {
    in = read_msh2(path);
  } catch (const std::exception &ex) {
    set_error(ex.what());
    return BES_ERR_MESH;
  }
  return create_from_input(in, geom_mode, gas, out);
}

void bes_implicit_free(bes_ctx *c);   // implicit.hip

int bes_ctx_destroy(bes_ctx *c) {
  if (!c) return BES_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  bes_implicit_free(c);
  void *ptrs[] = {c->dm.vol, c->dm.eps2, c->dm.sproj, c->dm.nbr, c->dm.Sg, c->dm.w, c->dm.rl, c->dm.nf,
                  c->dm.af, c->dm.rs, c->dm.rn, c->dm.info, c->U0, c->Ua, c->Ub, c->LG, c->G, c->phi,
                  c->dt, c->dtg, c->R, c->red, c->err, c->linf};
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (auto &p : c->ev_pending) { c->ev_pool.push_back(p.a); c->ev_pool.push_back(p.b); }
  for (hipEvent_t e : c->ev_pool) if (e) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return BES_OK;
}

int bes_ctx_info(const bes_ctx *c, int *ndim, long *ncell, long *nface) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  if (ndim) *ndim = c->nd;
  if (ncell) *ncell = (long)c->geo.ncell;
  if (nface) *nface = (long)c->geo.nface;
  return BES_OK;
}

int bes_get_geometry(const bes_ctx *c, long *owner, long *nbr, int *btype, double *S, double *n,
                     double *area, double *vol, double *centroid, long *cell_faces, long *cell_adj,
                     double *w, double *r) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  const Geometry &g = c->geo;
  if (owner) std::copy(g.owner.begin(), g.owner.end(), owner);
  if (nbr) std::copy(g.nbr.begin(), g.nbr.end(), nbr);
  if (btype) std::copy(g.btype.begin(), g.btype.end(), btype);
  if (S) std::copy(g.S.begin(), g.S.end(), S);
  if (n) std::copy(g.n.begin(), g.n.end(), n);
  if (area) std::copy(g.area.begin(), g.area.end(), area);
  if (vol) std::copy(g.vol.begin(), g.vol.end(), vol);
  if (centroid) std::copy(g.centroid.begin(), g.centroid.end(), centroid);
  if (cell_faces) std::copy(g.cface.begin(), g.cface.end(), cell_faces);
  if (cell_adj) std::copy(g.cadj.begin(), g.cadj.end(), cell_adj);
  if (w) std::copy(g.w.begin(), g.w.end(), w);
  if (r) std::copy(g.r.begin(), g.r.end(), r);
  return BES_OK;
}

int bes_set_freestream(bes_ctx *c, const double *prim) {
  if (!c || !prim) { set_error("null argument"); return BES_ERR_ARG; }
  for (int q = 0; q < c->nv + 1; q++) c->finf[q] = prim[q];
  return BES_OK;
}

int bes_set_state(bes_ctx *c, const double *U) {
  if (!c || !U) { set_error("null argument"); return BES_ERR_ARG; }
  return state_to_device(c, U, c->nv, c->U0);
}

int bes_get_state(bes_ctx *c, double *U) {
  if (!c || !U) { set_error("null argument"); return BES_ERR_ARG; }
  return state_to_host(c, c->U0, c->nv, U);
}

int bes_compute_time_step(bes_ctx *c, double cfl, int local) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  int s = launch_timestep(c, c->U0, cfl, local);
  if (s) return s;
  c->dt_local = local != 0;
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

// TimeStepCalculator::getTimeStep for every element (the global minimum on
// every element after a global computation, TimeStepCalculatorGlobal.cpp:100-102)
int bes_get_time_step(bes_ctx *c, double *dt) {
  if (!c || !dt) { set_error("null argument"); return BES_ERR_ARG; }
  if (c->dt_local) return state_to_host(c, c->dt, 1, dt);
  double g = 0;
  BES_HIP(hipMemcpy(&g, c->dtg, sizeof(double), hipMemcpyDeviceToHost));
  std::fill(dt, dt + c->N, g);
  return BES_OK;
}

int bes_compute_gradient_limiter(bes_ctx *c) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  const size_t N = c->N;
  if (!c->G) {
    int s;
    if ((s = dalloc(&c->G, (size_t)c->nv * c->nd * N)) || (s = dalloc(&c->phi, (size_t)c->nv * N))) return s;
  }
  int s = launch_gradlim(c, c->U0, c->G, c->phi);
  if (s) return s;
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

int bes_get_gradient(bes_ctx *c, double *G) {
  if (!c || !G || !c->G) { set_error("gradient not computed"); return BES_ERR_ARG; }
  // device [v][d][N] -> host [cell][v][d]
  return state_to_host(c, c->G, c->nv * c->nd, G);
}

int bes_get_limiter(bes_ctx *c, double *phi) {
  if (!c || !phi || !c->phi) { set_error("limiter not computed"); return BES_ERR_ARG; }
  return state_to_host(c, c->phi, c->nv, phi);
}

int bes_compute_residual(bes_ctx *c, double *R_out, double *maxabs_out) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  if (!c->R) { int s = dalloc(&c->R, (size_t)c->nv * c->N); if (s) return s; }
  int s;
  if ((s = reset_err(c))) return s;
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->red, 8, 0ull);
  if ((s = launch_residual(c, c->U0, nullptr, nullptr, 0.0, 1, c->R, c->red))) return s;
  if (maxabs_out) {
    unsigned long long b[8];
    BES_HIP(hipMemcpyAsync(b, c->red, sizeof(b), hipMemcpyDeviceToHost, c->stream));
    BES_HIP(hipStreamSynchronize(c->stream));
    for (int v = 0; v < c->nv; v++) { double x; std::memcpy(&x, &b[v], 8); maxabs_out[v] = x; }
  }
  if (R_out && (s = state_to_host(c, c->R, c->nv, R_out))) return s;
  return check_err(c);
}

int bes_explicit_run(bes_ctx *c, int niter, int nstage, const double *alpha, double cfl, int local,
                     double *linf_out) {
  if (!c || niter < 0 || nstage < 1 || nstage > 16 || !alpha) { set_error("invalid argument"); return BES_ERR_ARG; }
  const int nv = c->nv;
  if (linf_out && c->linf_cap < niter) {
    if (c->linf) hipFree(c->linf);
    c->linf = nullptr;
    c->linf_cap = 0;
    int s = dalloc(&c->linf, (size_t)niter * nv);
    if (s) return s;
    c->linf_cap = niter;
  }
  int s;
  if ((s = reset_err(c))) return s;
  unsigned long long *bits = c->red + 8;
  for (int it = 0; it < niter; it++) {
    if ((s = launch_timestep(c, c->U0, cfl, local))) return s;
    for (int st = 0; st < nstage; st++) {
      const bool last = st == nstage - 1;
      const double *src = st == 0 ? c->U0 : ((st - 1) % 2 == 0 ? c->Ua : c->Ub);
      double *dst = (last && nstage > 1) ? c->U0 : (st % 2 == 0 ? c->Ua : c->Ub);
      if ((s = launch_gradlim(c, src, nullptr, nullptr))) return s;
      if (last && linf_out) hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, bits, nv, 0ull);
      if ((s = launch_residual(c, src, c->U0, dst, alpha[st], local, nullptr, (last && linf_out) ? bits : nullptr)))
        return s;
      if (last && linf_out)
        hipLaunchKernelGGL(k_bits_to_double, dim3(1), dim3(64), 0, c->stream, bits, c->linf + (size_t)it * nv, nv);
      if (last && nstage == 1) std::swap(c->U0, c->Ua);
    }
  }
  if ((s = prof_collect(c))) return s;
  if (linf_out && niter > 0) {
    BES_HIP(hipMemcpyAsync(linf_out, c->linf, (size_t)niter * nv * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  BES_HIP(hipStreamSynchronize(c->stream));
  return check_err(c);
}

static int batch_common(int ndim, int kind, long n, const double *UL, const double *UR,
                        const double *nrm, const bes_gas *gas, double *out1, double *out2, bool jac) {
  if ((ndim != 2 && ndim != 3) || n < 0 || (kind < 1 || kind > 3) || (n > 0 && (!UL || !nrm || !out1)) ||
      (kind != kSlipWall && n > 0 && !UR)) {
    set_error("invalid batch argument");
    return BES_ERR_ARG;
  }
  if (n == 0) return BES_OK;
  const int nv = ndim + 2;
  GasConst gc = make_gas(gas);
  std::vector<double> ur((size_t)n * (nv + 1), 0.0);
  if (UR) std::copy(UR, UR + (size_t)n * (nv + 1), ur.begin());
  double *dUL, *dUR, *dn, *d1, *d2 = nullptr;
  const size_t o1 = jac ? (size_t)n * nv * nv : (size_t)n * nv;
  BES_HIP(hipMalloc(&dUL, (size_t)n * nv * 8));
  BES_HIP(hipMalloc(&dUR, (size_t)n * (nv + 1) * 8));
  BES_HIP(hipMalloc(&dn, (size_t)n * ndim * 8));
  BES_HIP(hipMalloc(&d1, o1 * 8));
  if (jac) BES_HIP(hipMalloc(&d2, o1 * 8));
  BES_HIP(hipMemcpy(dUL, UL, (size_t)n * nv * 8, hipMemcpyHostToDevice));
  BES_HIP(hipMemcpy(dUR, ur.data(), (size_t)n * (nv + 1) * 8, hipMemcpyHostToDevice));
  BES_HIP(hipMemcpy(dn, nrm, (size_t)n * ndim * 8, hipMemcpyHostToDevice));
  if (jac) BES_HIP(hipMemset(d2, 0, o1 * 8));
  const int bs = 128;
  if (!jac) {
    const int nb = grid_for((int)n, bs);
    if (ndim == 2) hipLaunchKernelGGL(k_flux_batch<2>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, d1);
    else hipLaunchKernelGGL(k_flux_batch<3>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, d1);
  } else {
    const int nb = grid_for((int)n * 2 * nv, bs);
    if (ndim == 2) hipLaunchKernelGGL(k_jacobian_batch<2>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, d1, d2);
    else hipLaunchKernelGGL(k_jacobian_batch<3>, dim3(nb), dim3(bs), 0, 0, kind, (int)n, dUL, dUR, dn, gc, d1, d2);
  }
  BES_HIP(hipGetLastError());
  BES_HIP(hipMemcpy(out1, d1, o1 * 8, hipMemcpyDeviceToHost));
  if (jac && out2) BES_HIP(hipMemcpy(out2, d2, o1 * 8, hipMemcpyDeviceToHost));
  hipFree(dUL); hipFree(dUR); hipFree(dn); hipFree(d1);
  if (d2) hipFree(d2);
  return BES_OK;
}

int bes_flux_batch(int ndim, int kind, long n, const double *UL, const double *UR, const double *normals,
                   const bes_gas *gas, double *F) {
  return batch_common(ndim, kind, n, UL, UR, normals, gas, F, nullptr, false);
}

int bes_jacobian_batch(int ndim, int kind, long n, const double *UL, const double *UR,
                       const double *normals, const bes_gas *gas, double *JL, double *JR) {
  return batch_common(ndim, kind, n, UL, UR, normals, gas, JL, JR, true);
}

long bes_last_error_face(const bes_ctx *c) { return c ? c->last_err_face : -1; }

int bes_set_profiling(bes_ctx *c, int on) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  int s = prof_collect(c);
  if (s) return s;
  c->prof = on != 0;
  for (int q = 0; q < bes_ctx::kProfKinds; q++) { c->prof_ms[q] = 0; c->prof_n[q] = 0; c->prof_bytes[q] = 0; }
  return BES_OK;
}

int bes_get_profile(bes_ctx *c, double *total_ms, long *launches, double *bytes, int nkinds) {
  if (!c || nkinds < 0) { set_error("invalid argument"); return BES_ERR_ARG; }
  int s = prof_collect(c);
  if (s) return s;
  for (int q = 0; q < nkinds && q < bes_ctx::kProfKinds; q++) {
    if (total_ms) total_ms[q] = c->prof_ms[q];
    if (launches) launches[q] = c->prof_n[q];
    if (bytes) bytes[q] = c->prof_bytes[q];
  }
  return BES_OK;
}

int bes_synchronize(bes_ctx *c) {
  if (!c) { set_error("null context"); return BES_ERR_ARG; }
  BES_HIP(hipStreamSynchronize(c->stream));
  return BES_OK;
}

}  // extern "C"
