// MI355X (gfx950) implementation of the C-ABI in include/fvs2d.h.
//
// One residual evaluation is three kinds of launch:
//   k_grad_*   cell-parallel: gradient of the primitive state (LSQ, GGCB or
//              GGNB) and, fused into the same pass over the stencil, the
//              min/max and face extrapolation of the limiter;
//   k_flux_*   face-parallel: MUSCL/UMUSCL reconstruction + Roe flux per
//              face, written once as flux*area (no float atomics);
//   k_update   cell-parallel: gathers its faces' fluxes in the reference's
//              per-cell order (a fixed gather table), forms -R/vol and applies
//              the Runge-Kutta stage update, the local time step, the
//              primitive state of the new stage and the residual-norm partial
//              sums in the same pass.
// Reductions are two-level and in a fixed order, so results are bitwise
// reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <numeric>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/fvs2d.h"
#include "kernels.hpp"
#include "mesh_prep.hpp"

using namespace fvs;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPC(x)                                                                           \
  do {                                                                                    \
    hipError_t e_ = (x);                                                                  \
    if (e_ != hipSuccess) return fail(FVS_EHIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int kBlock = 256;
constexpr int kSmallBlock = 128;  // the other block size the fused flux kernel can run with
constexpr int kVerr = 15;  // vortex-error record: 13 norms + location of the largest density error
constexpr int kSlice = 64;  // one wavefront

__host__ __device__ inline int nblk(long n, int b = kBlock) { return int((n + b - 1) / b); }

// XCD-aware block numbering: MI355X dispatches consecutive workgroups to its
// 8 XCDs round robin; remapping gives each XCD one contiguous eighth of the
// index range, so neighbour gathers stay in that XCD's L2.  Launch
// xcd_grid(nb) workgroups; logical blocks >= nb exit.
constexpr int kXcd = 8;
// A negative nb passed to a kernel means "plain numbering" (the autotuner
// picks per kernel class which one is faster on the device).
__device__ __forceinline__ int xcd_block(int nb) {
  if (nb < 0) return int(blockIdx.x);
  const int per = (nb + kXcd - 1) / kXcd;
  return int(blockIdx.x % kXcd) * per + int(blockIdx.x / kXcd);
}
inline int xcd_grid(int nb) { return ((nb + kXcd - 1) / kXcd) * kXcd; }
inline int xcd_arg(int nb, bool on) { return on ? nb : -nb; }

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ void ld4(const double *__restrict__ p, long i, double v[4]) {
  const double2 a = reinterpret_cast<const double2 *>(p)[2 * i];
  const double2 b = reinterpret_cast<const double2 *>(p)[2 * i + 1];
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
__device__ __forceinline__ void st4(double *__restrict__ p, long i, const double v[4]) {
  reinterpret_cast<double2 *>(p)[2 * i] = make_double2(v[0], v[1]);
  reinterpret_cast<double2 *>(p)[2 * i + 1] = make_double2(v[2], v[3]);
}
// a loaded state made primitive: with cons it holds conservative variables
// and is converted in place (called after the loads that travel with it are
// issued, so the conversion waits on the state alone)
__device__ __forceinline__ void make_prim(bool cons, double g, double w[4]) {
  if (cons) {
    const double u[4] = {w[0], w[1], w[2], w[3]};
    cons2prim(u, g, w);
  }
}
__device__ __forceinline__ void ld8(const double *__restrict__ p, long i, double v[8]) {
  const double2 *q = reinterpret_cast<const double2 *>(p) + 4 * i;
  double2 a = q[0], b = q[1], c = q[2], d = q[3];
  v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = c.x; v[5] = c.y; v[6] = d.x; v[7] = d.y;
}
__device__ __forceinline__ void st8(double *__restrict__ p, long i, const double v[8]) {
  double2 *q = reinterpret_cast<double2 *>(p) + 4 * i;
  q[0] = make_double2(v[0], v[1]); q[1] = make_double2(v[2], v[3]);
  q[2] = make_double2(v[4], v[5]); q[3] = make_double2(v[6], v[7]);
}

__global__ void k_prim(int N, double g, const double *__restrict__ U, double *__restrict__ W) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double u[4], w[4];
  ld4(U, i, u);
  cons2prim(u, g, w);
  st4(W, i, w);
}

// Limiter of one cell from its stencil min/max and its gradient
// (src/gradient_limiter.f90:47-91): phi = min over variables of
// min(1, min over faces of L(dmax|dmin, q_f - q)).  For Venkatakrishnan and
// Barth every candidate L = n/d has d > 0 (Venkatakrishnan's denominator is
// (a + b/2)^2 + 7b^2/4 + eps^2; Barth's quotient a/b is taken with both
// signs flipped when b < 0), so the smallest one is found by comparing
// n_i d_j with n_j d_i and divided out once per cell instead of once per
// face and variable; the result is the selected candidate's own quotient,
// which can differ from the reference's minimum only where two candidates
// are within rounding of each other.  Van Albada divides per face as
// written.
__device__ __forceinline__ double cell_limiter(int lt, int nf, int N, int i, const double2 *__restrict__ off,
                                               double vol, const double w[4], const double mn[4],
                                               const double mx[4], const double gx[4], const double gy[4]) {
  // face offsets in registers: every index below is a compile-time constant
  double2 o[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) o[k] = (k < nf) ? off[long(k) * N + i] : make_double2(0.0, 0.0);
  if (lt == 3) {
    double phi = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double q = w[v], dmax = mx[v] - q, dmin = mn[v] - q;
      double pe = 100.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        if (k >= nf) break;
        const double pf = q + o[k].x * gx[v] + o[k].y * gy[v];
        const double diff = pf - q;
        double e = 1.0;
        if (diff > 0.0) e = limiter_fn(lt, dmax, diff, vol);
        else if (diff < 0.0) e = limiter_fn(lt, dmin, diff, vol);
        pe = (e < pe) ? e : pe;
      }
      const double pv = (1.0 < pe) ? 1.0 : pe;
      phi = (v == 0 || pv < phi) ? pv : phi;
    }
    return phi;
  }
  const double eps2 = venkat_eps2(vol);
  double bn = 1.0, bd = 1.0;  // smallest candidate so far (the cap at 1 to start)
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double q = w[v], dmax = mx[v] - q, dmin = mn[v] - q;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k >= nf) break;
      const double pf = q + o[k].x * gx[v] + o[k].y * gy[v];
      const double b = pf - q;
      if (b != 0.0) {
        const double a = (b > 0.0) ? dmax : dmin;
        double n, d;
        if (lt == 1) {
          n = (a * a + eps2) + 2.0 * b * a;
          d = a * a + 2.0 * (b * b) + a * b + eps2;
        } else {
          n = (b > 0.0) ? a : -a;
          d = (b > 0.0) ? b : -b;
        }
        if (n * bd < bn * d) {
          bn = n;
          bd = d;
        }
      }
    }
  }
  return bn / bd;
}

// LSQ gradient (src/gradient_lsq.f90:371-419): grad_i = sum_j cw_j (q_j - q_i),
// cw = G^-1 D^T w per stencil entry; LIM fuses the limiter.
template <int LIM>
__global__ __launch_bounds__(kBlock) void k_grad_lsq(int nbx, int N, const double *__restrict__ W, const int *__restrict__ soff,
                                                     const int *__restrict__ swid, const int *__restrict__ snb,
                                                     const double2 *__restrict__ scw, double *__restrict__ grad,
                                                     double *__restrict__ phi, int lt, const int *__restrict__ nvrt,
                                                     const double2 *__restrict__ loff, const double *__restrict__ vol) {
  const int i = xcd_block(nbx) * kBlock + threadIdx.x;
  if (i >= N) return;
  const int s = i / kSlice, l = i % kSlice;
  const int base = soff[s], wid = swid[s];
  double w[4], gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 0, 0}, mn[4], mx[4];
  ld4(W, i, w);
  for (int v = 0; v < 4; ++v) mn[v] = mx[v] = w[v];
  for (int k = 0; k < wid; ++k) {
    const long e = base + long(k) * kSlice + l;
    const int j = snb[e];
    const double2 c = scw[e];
    double q[4];
    ld4(W, j, q);
    for (int v = 0; v < 4; ++v) {
      const double d = q[v] - w[v];
      gx[v] = gx[v] + c.x * d;
      gy[v] = gy[v] + c.y * d;
      if (LIM) {
        mn[v] = (q[v] < mn[v]) ? q[v] : mn[v];
        mx[v] = (q[v] > mx[v]) ? q[v] : mx[v];
      }
    }
  }
  double g8[8] = {gx[0], gx[1], gx[2], gx[3], gy[0], gy[1], gy[2], gy[3]};
  st8(grad, i, g8);
  if (LIM) phi[i] = cell_limiter(lt, nvrt[i], N, i, loff, vol[i], w, mn, mx, gx, gy);
}

// stencil min/max over the limiter's (node-neighbour) stencil, for GG gradients
__device__ __forceinline__ void stencil_minmax(int i, const int *__restrict__ soff, const int *__restrict__ swid,
                                               const int *__restrict__ snb, const double *__restrict__ W,
                                               const double w[4], double mn[4], double mx[4]) {
  const int s = i / kSlice, l = i % kSlice;
  for (int v = 0; v < 4; ++v) mn[v] = mx[v] = w[v];
  const int wid = swid[s];
  const long base = soff[s] + l;
  // four entries per round: their ids, then their states, all in flight together
  for (int k0 = 0; k0 < wid; k0 += 4) {
    int j[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) j[u] = (k0 + u < wid) ? snb[base + long(k0 + u) * kSlice] : i;
    double q[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) ld4(W, j[u], q[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      for (int v = 0; v < 4; ++v) {
        mn[v] = (q[u][v] < mn[v]) ? q[u][v] : mn[v];
        mx[v] = (q[u][v] > mx[v]) ? q[u][v] : mx[v];
      }
  }
}

// GGCB gradient (src/gradient_ggcb.f90:116-138):
// grad_i = (coef0 q_i + sum_f coefnb_f q_nb(f)) / vol
template <int LIM>
__global__ __launch_bounds__(kBlock) void k_grad_ggcb(int nbx, int N, int nfw, const double *__restrict__ W,
                                                      int cons, double gamma,
                                                      const double2 *__restrict__ c0, const int *__restrict__ nb,
                                                      const double2 *__restrict__ cn, const double *__restrict__ vol,
                                                      double *__restrict__ grad, double *__restrict__ phi, int lt,
                                                      const int *__restrict__ nvrt, const double2 *__restrict__ loff,
                                                      const int *__restrict__ soff, const int *__restrict__ swid,
                                                      const int *__restrict__ snb) {
  const int i = xcd_block(nbx) * kBlock + threadIdx.x;
  if (i >= N) return;
  // every face neighbour's id and coefficient, then every state, in flight
  // together (nfw <= 4); the sum keeps the face order
  int j[4];
  double2 c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    j[k] = (k < nfw) ? nb[long(k) * N + i] : i;
    c[k] = (k < nfw) ? cn[long(k) * N + i] : make_double2(0.0, 0.0);
  }
  double w[4], gx[4], gy[4], q[4][4];
  ld4(W, i, w);
  const double2 a = c0[i];
  const double V = vol[i];
#pragma unroll
  for (int k = 0; k < 4; ++k) ld4(W, j[k], q[k]);
  // without a limiter W may hold conservative states (SSP-RK3 stages)
  const bool cv = !LIM && cons;
  make_prim(cv, gamma, w);
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < nfw) make_prim(cv, gamma, q[k]);
  for (int v = 0; v < 4; ++v) gx[v] = a.x * w[v], gy[v] = a.y * w[v];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < nfw)
      for (int v = 0; v < 4; ++v) {
        gx[v] = gx[v] + c[k].x * q[k][v];
        gy[v] = gy[v] + c[k].y * q[k][v];
      }
  for (int v = 0; v < 4; ++v) gx[v] = gx[v] / V, gy[v] = gy[v] / V;
  double g8[8] = {gx[0], gx[1], gx[2], gx[3], gy[0], gy[1], gy[2], gy[3]};
  st8(grad, i, g8);
  if (LIM) {
    double mn[4], mx[4];
    stencil_minmax(i, soff, swid, snb, W, w, mn, mx);
    phi[i] = cell_limiter(lt, nvrt[i], N, i, loff, V, w, mn, mx, gx, gy);
  }
}

// Limiter bounds of a Green-Gauss gradient from a staged halo: the block's
// own states (local 0..255) and the outside cells of its node-neighbour
// stencils (from 256) are staged in LDS, and each cell walks its stencil as
// 16-bit local codes (two per 32-bit word, the LSQ staged-halo layout) in LDS
// instead of gathering the states from HBM.  min and max are exact, so the
// bounds are those of the stencil walk in HBM (stencil_minmax).  With cons, W
// holds conservative states (SSP-RK3 stages), converted as they are staged.
__device__ __forceinline__ void gg_stage_halo(int lb, int N, const double *__restrict__ W, int cons, double gamma,
                                              const int *__restrict__ ghoff, const int *__restrict__ ghlist,
                                              double *__restrict__ lW) {
  const int i0 = lb * kBlock, id = i0 + threadIdx.x;
  if (id < N) {
    double w[4];
    ld4(W, id, w);
    make_prim(cons, gamma, w);
    st4(lW, threadIdx.x, w);
  }
  const int o2 = ghoff[lb], nh = ghoff[lb + 1] - o2;
  for (int q = threadIdx.x; q < nh; q += kBlock) {
    double w[4];
    ld4(W, ghlist[o2 + q], w);
    make_prim(cons, gamma, w);
    st4(lW, kBlock + q, w);
  }
}
__device__ __forceinline__ void gg_halo_minmax(const unsigned *__restrict__ lc, int wid,
                                               const double *__restrict__ lW, const double w[4], double mn[4],
                                               double mx[4]) {
  for (int v = 0; v < 4; ++v) mn[v] = mx[v] = w[v];
  for (int k = 0; k < wid; ++k) {
    const int j = int(((lc[(k >> 1) * kSlice] >> (16 * (k & 1))) & 0xffffu) >> 4);
    double q[4];
    ld4(lW, j, q);
    for (int v = 0; v < 4; ++v) {
      mn[v] = (q[v] < mn[v]) ? q[v] : mn[v];
      mx[v] = (q[v] > mx[v]) ? q[v] : mx[v];
    }
  }
}

// GGCB gradient with the limiter, bounds from the staged halo (one block of
// 256 cells per workgroup; LDS 32 B per local cell).  With cons, W holds
// conservative states: the staged ones and the face neighbours are converted
// after their loads are issued.
__global__ __launch_bounds__(kBlock) void k_grad_ggcb_h(int nbx, int N, int nfw, const double *__restrict__ W,
                                                        int cons, double gamma,
                                                        const double2 *__restrict__ c0, const int *__restrict__ nb,
                                                        const double2 *__restrict__ cn, const double *__restrict__ vol,
                                                        double *__restrict__ grad, double *__restrict__ phi, int lt,
                                                        const int *__restrict__ nvrt, const double2 *__restrict__ loff,
                                                        const int *__restrict__ ghoff, const int *__restrict__ ghlist,
                                                        const int *__restrict__ poff, const int *__restrict__ swid,
                                                        const unsigned *__restrict__ lcode) {
  extern __shared__ double lds[];
  const int lb = xcd_block(nbx);
  if (lb >= (nbx < 0 ? -nbx : nbx)) return;  // the whole workgroup leaves before the barrier
  const int i = lb * kBlock + threadIdx.x;
  const bool act = i < N;
  // the gradient's global loads are in flight while the block is staged
  int j[4];
  double2 c[4];
  double w[4], q[4][4];
  double2 a = make_double2(0.0, 0.0);
  double V = 1.0;
  if (act) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      j[k] = (k < nfw) ? nb[long(k) * N + i] : i;
      c[k] = (k < nfw) ? cn[long(k) * N + i] : make_double2(0.0, 0.0);
    }
    ld4(W, i, w);
    a = c0[i];
    V = vol[i];
#pragma unroll
    for (int k = 0; k < 4; ++k) ld4(W, j[k], q[k]);
  }
  gg_stage_halo(lb, N, W, cons, gamma, ghoff, ghlist, lds);
  if (act) {
    make_prim(cons, gamma, w);
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < nfw) make_prim(cons, gamma, q[k]);
  }
  __syncthreads();
  if (!act) return;
  double gx[4], gy[4];
  for (int v = 0; v < 4; ++v) gx[v] = a.x * w[v], gy[v] = a.y * w[v];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < nfw)
      for (int v = 0; v < 4; ++v) {
        gx[v] = gx[v] + c[k].x * q[k][v];
        gy[v] = gy[v] + c[k].y * q[k][v];
      }
  for (int v = 0; v < 4; ++v) gx[v] = gx[v] / V, gy[v] = gy[v] / V;
  double g8[8] = {gx[0], gx[1], gx[2], gx[3], gy[0], gy[1], gy[2], gy[3]};
  st8(grad, i, g8);
  const int s = i / kSlice;
  double mn[4], mx[4];
  gg_halo_minmax(lcode + poff[s] + threadIdx.x % kSlice, swid[s], lds, w, mn, mx);
  phi[i] = cell_limiter(lt, nvrt[i], N, i, loff, V, w, mn, mx, gx, gy);
}

// GGNB gradient (src/gradient_ggnb.f90:351-378) in two passes.  The
// reference's per-cell sum coef0 q_i + sum_nodes coefedg(in) sum_j coefnb q_j
// regroups by node into sum_nodes coefedg(in) F(node), F(v) = sum over the
// cells j around v of idw(v)/d(j, v) q_j: an inverse-distance average
// formed once per node, 12 B per incidence, instead of 20 B per cell-node-
// cell incidence.  Agrees with the reference to rounding (the grouping).
//
// Pass 1: node values, node incidences in 64-node slices (entry k of lane l
// of slice s at nsoff[s] + 64 k + l; padding: a cell of the node, weight 0).
__global__ __launch_bounds__(kBlock) void k_ggnb_nodes(int NV, const double *__restrict__ W, int cons,
                                                       double gamma, const int *__restrict__ nsoff,
                                                       const int *__restrict__ nswid,
                                                       const int *__restrict__ nscell,
                                                       const double *__restrict__ nscoef, double *__restrict__ F) {
  const int v = blockIdx.x * kBlock + threadIdx.x;
  if (v >= NV) return;
  const int s = v / kSlice, l = v % kSlice;
  const int base = nsoff[s], wid = nswid[s];
  double f[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < wid; k0 += 4) {
    int j[4];
    double c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long e = base + long(k0 + u) * kSlice + l;
      j[u] = (k0 + u < wid) ? nscell[e] : nscell[base + l];
      c[u] = (k0 + u < wid) ? nscoef[e] : 0.0;
    }
    double q[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) ld4(W, j[u], q[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) make_prim(cons, gamma, q[u]);  // conservative states (SSP-RK3 stages)
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (k0 + u < wid)
        for (int v2 = 0; v2 < 4; ++v2) f[v2] = f[v2] + c[u] * q[u][v2];
  }
  st4(F, v, f);
}

// Pass 2: grad_i = sum over the cell's nodes (gw slots, planes [k][N]; a
// padded slot has weight 0) of coefedg * F(node), / vol; with LIM the limiter
// over the node-neighbour stencil.
template <int LIM>
__global__ __launch_bounds__(kBlock) void k_grad_ggnb(int nbx, int N, int gw, const double *__restrict__ W,
                                                      const int *__restrict__ cnode, const double2 *__restrict__ cedg,
                                                      const double *__restrict__ F, const double *__restrict__ vol,
                                                      double *__restrict__ grad, double *__restrict__ phi, int lt,
                                                      const int *__restrict__ nvrt, const double2 *__restrict__ loff,
                                                      const int *__restrict__ soff, const int *__restrict__ swid,
                                                      const int *__restrict__ snb) {
  const int i = xcd_block(nbx) * kBlock + threadIdx.x;
  if (i >= N) return;
  int nd[4];
  double2 c[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    nd[k] = (k < gw) ? cnode[long(k) * N + i] : 0;
    c[k] = (k < gw) ? cedg[long(k) * N + i] : make_double2(0.0, 0.0);
  }
  double f[4][4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < gw) ld4(F, nd[k], f[k]);
  double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < gw)
      for (int v = 0; v < 4; ++v) {
        gx[v] = gx[v] + c[k].x * f[k][v];
        gy[v] = gy[v] + c[k].y * f[k][v];
      }
  const double V = vol[i];
  for (int v = 0; v < 4; ++v) gx[v] = gx[v] / V, gy[v] = gy[v] / V;
  double g8[8] = {gx[0], gx[1], gx[2], gx[3], gy[0], gy[1], gy[2], gy[3]};
  st8(grad, i, g8);
  if (LIM) {
    double w[4], mn[4], mx[4];
    ld4(W, i, w);
    stencil_minmax(i, soff, swid, snb, W, w, mn, mx);
    phi[i] = cell_limiter(lt, nvrt[i], N, i, loff, V, w, mn, mx, gx, gy);
  }
}

// GGNB pass 2 with the limiter, bounds from the staged halo
__global__ __launch_bounds__(kBlock) void k_grad_ggnb_h(int nbx, int N, int gw, const double *__restrict__ W,
                                                        int cons, double gamma, const int *__restrict__ cnode, const double2 *__restrict__ cedg,
                                                        const double *__restrict__ F, const double *__restrict__ vol,
                                                        double *__restrict__ grad, double *__restrict__ phi, int lt,
                                                        const int *__restrict__ nvrt, const double2 *__restrict__ loff,
                                                        const int *__restrict__ ghoff, const int *__restrict__ ghlist,
                                                        const int *__restrict__ poff, const int *__restrict__ swid,
                                                        const unsigned *__restrict__ lcode) {
  extern __shared__ double lds[];
  const int lb = xcd_block(nbx);
  if (lb >= (nbx < 0 ? -nbx : nbx)) return;  // the whole workgroup leaves before the barrier
  const int i = lb * kBlock + threadIdx.x;
  const bool act = i < N;
  int nd[4];
  double2 c[4];
  double f[4][4], w[4];
  double V = 1.0;
  if (act) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      nd[k] = (k < gw) ? cnode[long(k) * N + i] : 0;
      c[k] = (k < gw) ? cedg[long(k) * N + i] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < gw) ld4(F, nd[k], f[k]);
    ld4(W, i, w);
    V = vol[i];
  }
  gg_stage_halo(lb, N, W, cons, gamma, ghoff, ghlist, lds);
  if (act) make_prim(cons, gamma, w);
  __syncthreads();
  if (!act) return;
  double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (k < gw)
      for (int v = 0; v < 4; ++v) {
        gx[v] = gx[v] + c[k].x * f[k][v];
        gy[v] = gy[v] + c[k].y * f[k][v];
      }
  for (int v = 0; v < 4; ++v) gx[v] = gx[v] / V, gy[v] = gy[v] / V;
  double g8[8] = {gx[0], gx[1], gx[2], gx[3], gy[0], gy[1], gy[2], gy[3]};
  st8(grad, i, g8);
  const int s = i / kSlice;
  double mn[4], mx[4];
  gg_halo_minmax(lcode + poff[s] + threadIdx.x % kSlice, swid[s], lds, w, mn, mx);
  phi[i] = cell_limiter(lt, nvrt[i], N, i, loff, V, w, mn, mx, gx, gy);
}

// interior faces: reconstruction (src/residual.f90:87-92) + Roe (95)
__global__ __launch_bounds__(kBlock) void k_flux_int(int nbx, int Fi, Phys P, const int2 *__restrict__ fc,
                                                     const double *__restrict__ geo, const double *__restrict__ W,
                                                     const double *__restrict__ grad, const double *__restrict__ phi,
                                                     double *__restrict__ FA, double *__restrict__ WSA) {
  const int f = xcd_block(nbx) * kBlock + threadIdx.x;
  if (f >= Fi) return;
  const int2 c = fc[f];
  double g[8];
  ld8(geo, f, g);
  double wL[4], wR[4], gL[8], gR[8];
  ld4(W, c.x, wL);
  ld4(W, c.y, wR);
  ld8(grad, c.x, gL);
  ld8(grad, c.y, gR);
  const double phL = phi[c.x], phR = phi[c.y];
  const double k2 = P.kappa / 2.0, k1 = 1.0 - P.kappa;
  double pL[4], pR[4];
  for (int v = 0; v < 4; ++v) {
    const double gC = wR[v] - wL[v];
    const double dL = fma(g[3], gL[v], g[4] * gL[4 + v]);  // the FMAs of interior_flux_v
    const double dR = fma(g[5], gR[v], g[6] * gR[4 + v]);
    pL[v] = fma(phL, fma(k2, gC, k1 * dL), wL[v]);
    pR[v] = fma(phR, fma(-k2, gC, k1 * dR), wR[v]);
  }
  double fl[4], ws;
  roe_flux(P.gamma, pL, pR, g[0], g[1], fl, ws);
  const double A = g[2];
  double fa[4] = {fl[0] * A, fl[1] * A, fl[2] * A, fl[3] * A};
  st4(FA, f, fa);
  WSA[f] = ws * A;
}

// boundary faces: extrapolation (src/residual.f90:130), ghost state
// (bc_flux, 183-255), Roe (142)
__global__ __launch_bounds__(kBlock) void k_flux_bnd(int Fb, int Fi, Phys P, double t, const int *__restrict__ bcell,
                                                     const int *__restrict__ btype, const double *__restrict__ geo,
                                                     const double *__restrict__ W, const double *__restrict__ grad,
                                                     const double *__restrict__ phi, double *__restrict__ FA,
                                                     double *__restrict__ WSA, const int *__restrict__ binst,
                                                     const double *__restrict__ pinf_inst) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= Fb) return;
  const int c = bcell[b];
  double g[8], w[4], gr[8];
  ld8(geo, b, g);
  ld4(W, c, w);
  ld8(grad, c, gr);
  const double ph = phi[c];
  double pL[4], pR[4];
  for (int v = 0; v < 4; ++v) pL[v] = fma(ph, fma(g[3], gr[v], g[4] * gr[4 + v]), w[v]);
  const int ty = btype[b];
  const double nx = g[0], ny = g[1];
  if (ty == FVS_BC_FREESTREAM) {
    const double *pinf = binst ? pinf_inst + 4 * binst[b] : P.pinf;
    for (int v = 0; v < 4; ++v) pR[v] = pinf[v];
  } else if (ty == FVS_BC_SLIP_WALL) {
    const double un = pL[1] * nx + pL[2] * ny;
    pR[0] = pL[0];
    pR[1] = pL[1] - 2.0 * un * nx;
    pR[2] = pL[2] - 2.0 * un * ny;
    pR[3] = pL[3];
  } else if (P.lvortex) {
    vortex_state(P, t, g[5], g[6], pR);
  } else {
    mms_state(g[5], g[6], pR);
  }
  double fl[4], ws;
  roe_flux(P.gamma, pL, pR, nx, ny, fl, ws);
  const double A = g[2];
  double fa[4] = {fl[0] * A, fl[1] * A, fl[2] * A, fl[3] * A};
  st4(FA, Fi + b, fa);
  WSA[Fi + b] = ws * A;
}

enum { M_RESID = 0, M_SSP3 = 1, M_SSP42 = 2, M_CLASSIC = 3 };

struct Upd {
  int mode, steady, first, last;
  double a, b;   // SSP3: U = a*U0 + b*(U + h R)
  double h;      // unsteady step factor, or the stage constant when steady
  double coef;   // rk coefficient (4-stage schemes)
  int use_acc;   // CLASSIC final stage: update with the accumulator
  double cfl, gamma;
};

__device__ __forceinline__ void block_sum4(double v[4], double *red) {
  // fixed-shape tree over the block: deterministic
  const int t = threadIdx.x;
  for (int k = 0; k < 4; ++k) red[k * kBlock + t] = v[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < 4; ++k) red[k * kBlock + t] = red[k * kBlock + t] + red[k * kBlock + t + s];
    __syncthreads();
  }
  for (int k = 0; k < 4; ++k) v[k] = red[k * kBlock];
}

// gather (src/residual.f90:97-101,153-155 in the reference's per-cell order),
// -R/vol (164-166), local time step (src/runge_kutta.f90:433) and stage update
__global__ __launch_bounds__(kBlock) void k_update(int nbx, int N, int gw, Upd A, const int *__restrict__ gat,
                                                   const double *__restrict__ FA, const double *__restrict__ WSA,
                                                   const double *__restrict__ vol, double *__restrict__ Ua,
                                                   double *__restrict__ Ub, double *__restrict__ Facc,
                                                   double *__restrict__ W, double *__restrict__ dtl,
                                                   double *__restrict__ rout, double *__restrict__ wsout,
                                                   double *__restrict__ part) {
  __shared__ double red[4 * kBlock];
  const int nb = nblk(N), lb = xcd_block(nbx);
  if (lb >= nb) return;
  const int i = lb * kBlock + threadIdx.x;
  double d2[4] = {0, 0, 0, 0};
  if (i < N) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0}, wacc = 0.0;
    for (int k = 0; k < gw; ++k) {
      const int s = gat[long(k) * N + i];
      if (s == 0) break;
      const int slot = (s > 0 ? s : -s) - 1;
      double fa[4];
      ld4(FA, slot, fa);
      if (s > 0)
        for (int v = 0; v < 4; ++v) acc[v] = acc[v] + fa[v];
      else
        for (int v = 0; v < 4; ++v) acc[v] = acc[v] - fa[v];
      wacc = wacc + WSA[slot];
    }
    const double V = vol[i];
    double R[4];
    const double iV = 1.0 / V;  // one division per cell (src/residual.f90:164-166 divides each)
    for (int v = 0; v < 4; ++v) R[v] = -acc[v] * iV;
    if (A.mode == M_RESID) {
      st4(rout, i, R);
      wsout[i] = wacc;
    } else {
      double hs = A.h;
      if (A.steady) {
        double dl;
        if (A.first) {
          dl = A.cfl * V / (0.5 * wacc);
          dtl[i] = dl;
        } else {
          dl = dtl[i];
        }
        hs = dl * A.h;
      }
      double u0[4], un[4];
      ld4(Ua, i, u0);
      if (A.mode == M_SSP3) {
        double uc[4];
        if (A.first) for (int v = 0; v < 4; ++v) uc[v] = u0[v];
        else ld4(Ub, i, uc);
        for (int v = 0; v < 4; ++v) un[v] = A.a * u0[v] + A.b * (uc[v] + hs * R[v]);
      } else if (A.mode == M_SSP42) {
        double fo[4] = {0, 0, 0, 0};
        if (!A.first) ld4(Facc, i, fo);
        for (int v = 0; v < 4; ++v) un[v] = u0[v] + hs * (A.coef * R[v] + fo[v]);
        if (!A.last) {
          double fn[4];
          for (int v = 0; v < 4; ++v) fn[v] = fo[v] + R[v];
          st4(Facc, i, fn);
        }
      } else {
        double fo[4] = {0, 0, 0, 0}, fn[4];
        if (!A.first) ld4(Facc, i, fo);
        for (int v = 0; v < 4; ++v) fn[v] = fo[v] + A.coef * R[v];
        if (A.use_acc) for (int v = 0; v < 4; ++v) un[v] = u0[v] + hs * fn[v];
        else for (int v = 0; v < 4; ++v) un[v] = u0[v] + hs * R[v];
        if (!A.last) st4(Facc, i, fn);
      }
      if (A.last) {
        st4(Ua, i, un);
        for (int v = 0; v < 4; ++v) {
          const double d = fabs(un[v] - u0[v]);
          d2[v] = d * d;
        }
      } else if (A.mode == M_SSP3) {
        st4(Ub, i, un);
      }
      double w[4];
      cons2prim(un, A.gamma, w);
      st4(W, i, w);
    }
  }
  if (A.mode != M_RESID && A.last) {
    block_sum4(d2, red);
    if (threadIdx.x == 0)
      for (int v = 0; v < 4; ++v) part[4 * lb + v] = d2[v];
  }
}

// residual norms of one step: sqrt(sum_i (U-U0)^2 / N) per variable
// (src/runge_kutta.f90:169-184), partials summed in block order
constexpr int kNormThreads = 1024;
// (bump: also advance the step counter, which the caller then skips)
constexpr int kNormLoads = 8;
__global__ __launch_bounds__(kNormThreads) void k_norm_final(int nparts, int N, const double *__restrict__ part,
                                                             double *__restrict__ out, int *__restrict__ ctr,
                                                             int bump) {
  __shared__ double red[4 * kNormThreads];
  const int t = threadIdx.x;
  double s[4] = {0, 0, 0, 0};
  // thread t adds partials t, t + 1024, ... in that order; eight loads are in
  // flight before the adds (absent ones are exact zeros)
  for (int p0 = t; p0 < nparts; p0 += kNormLoads * kNormThreads) {
    double q[kNormLoads][4];
#pragma unroll
    for (int u = 0; u < kNormLoads; ++u) {
      const int p = p0 + u * kNormThreads;
      if (p < nparts) ld4(part, p, q[u]);
      else q[u][0] = q[u][1] = q[u][2] = q[u][3] = 0.0;
    }
#pragma unroll
    for (int u = 0; u < kNormLoads; ++u)
      for (int v = 0; v < 4; ++v) s[v] = s[v] + q[u][v];
  }
  for (int v = 0; v < 4; ++v) red[v * kNormThreads + t] = s[v];
  __syncthreads();
  for (int h = kNormThreads / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int v = 0; v < 4; ++v) red[v * kNormThreads + t] = red[v * kNormThreads + t] + red[v * kNormThreads + t + h];
    __syncthreads();
  }
  if (t == 0) {
    double *o = out + 4 * (ctr ? ctr[0] : 0);
    for (int v = 0; v < 4; ++v) o[v] = sqrt(red[v * kNormThreads] / double(N));
    if (bump && ctr) ctr[0] = ctr[0] + 1;
  }
}

// isentropic-vortex error over interior cells (src/mms.f90:315-361)
// error_isentropic_vortex (src/mms.f90:271-365) over the interior cells,
// listed in the reference's ascending order: per block the max, sum and sum
// of squares of |U - U_exact| per variable, and the list position of the
// largest density error (the first one on ties, like maxloc)
__global__ __launch_bounds__(kBlock) void k_vortex_err(int n, Phys P, double t, const int *__restrict__ cells,
                                                       const double *__restrict__ cx, const double *__restrict__ cy,
                                                       const double *__restrict__ U, double *__restrict__ part) {
  __shared__ double red[4 * kBlock];
  __shared__ int arg[kBlock];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double d[4] = {0, 0, 0, 0};
  if (k < n) {
    const int i = cells[k];
    double pv[4], ex[4], u[4];
    vortex_state(P, t, cx[i], cy[i], pv);
    ex[0] = pv[0];
    ex[1] = pv[0] * pv[1];
    ex[2] = pv[0] * pv[2];
    ex[3] = pv[3] / (P.gamma - 1.0) + 0.5 * ex[0] * (pv[1] * pv[1] + pv[2] * pv[2]);
    ld4(U, i, u);
    for (int v = 0; v < 4; ++v) d[v] = fabs(u[v] - ex[v]);
  }
  // max, and for the density the first position holding it
  double m[4] = {d[0], d[1], d[2], d[3]};
  int am = k < n ? k : n;
  {
    const int t2 = threadIdx.x;
    for (int q = 0; q < 4; ++q) red[q * kBlock + t2] = m[q];
    arg[t2] = am;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
      if (t2 < s) {
        const double a = red[t2], b = red[t2 + s];
        if (b > a || (b == a && arg[t2 + s] < arg[t2])) arg[t2] = arg[t2 + s];
        for (int q = 0; q < 4; ++q) red[q * kBlock + t2] = fmax(red[q * kBlock + t2], red[q * kBlock + t2 + s]);
      }
      __syncthreads();
    }
    for (int q = 0; q < 4; ++q) m[q] = red[q * kBlock];
    am = arg[0];
    __syncthreads();
  }
  double l1[4] = {d[0], d[1], d[2], d[3]};
  block_sum4(l1, red);
  __syncthreads();
  double l2[4] = {d[0] * d[0], d[1] * d[1], d[2] * d[2], d[3] * d[3]};
  block_sum4(l2, red);
  if (threadIdx.x == 0) {
    for (int q = 0; q < 4; ++q) {
      part[13 * blockIdx.x + q] = m[q];
      part[13 * blockIdx.x + 4 + q] = l1[q];
      part[13 * blockIdx.x + 8 + q] = l2[q];
    }
    part[13 * blockIdx.x + 12] = double(am);
  }
}

// the 13 norms of src/mms.f90:351-361 and the centre of the cell with the
// largest density error (src/mms.f90:363: maxloc over all cells of an array
// that is zero outside the interior cells, so cell `first` when every
// interior error is zero)
__global__ void k_vortex_final(int nparts, int n, const double *__restrict__ part, const int *__restrict__ cells,
                               const double *__restrict__ cx, const double *__restrict__ cy, int first,
                               double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double m[4] = {0, 0, 0, 0}, l1[4] = {0, 0, 0, 0}, l2[4] = {0, 0, 0, 0};
  int am = 0;
  for (int p = 0; p < nparts; ++p) {
    if (part[13 * p] > m[0]) am = int(part[13 * p + 12]);
    for (int q = 0; q < 4; ++q) {
      m[q] = fmax(m[q], part[13 * p + q]);
      l1[q] = l1[q] + part[13 * p + 4 + q];
      l2[q] = l2[q] + part[13 * p + 8 + q];
    }
  }
  const double dn = double(n);
  double tot = 0.0;
  for (int q = 0; q < 4; ++q) {
    out[3 * q] = m[q];
    out[3 * q + 1] = l1[q] / dn;
    out[3 * q + 2] = sqrt(l2[q] / dn);
    tot = tot + l2[q];
  }
  out[12] = sqrt(tot / dn);
  const int c = m[0] > 0.0 ? cells[am] : first;
  out[13] = cx[c];
  out[14] = cy[c];
}

// C_L / C_D over wall faces (src/io.f90:401-442), one thread in face order
__global__ void k_wall_forces(int nw, const int *__restrict__ wf, const int *__restrict__ bcell,
                              const double *__restrict__ geo, const double *__restrict__ W,
                              const double *__restrict__ grad, double mach, double pinf, double ca_, double sa_,
                              double *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double cn = 0, ca = 0, cn1 = 0, ca1 = 0;
  for (int k = 0; k < nw; ++k) {
    const int b = wf[k], c = bcell[b];
    double g[8], gr[8], w[4];
    ld8(geo, b, g);
    ld8(grad, c, gr);
    ld4(W, c, w);
    const double pw = w[3] + g[3] * gr[3] + g[4] * gr[7];
    const double cp = 2.0 / (mach * mach) * (pw - pinf);
    const double cp1 = 2.0 / (mach * mach) * (w[3] - pinf);
    cn = cn + cp * g[1] * g[2];
    ca = ca + cp * g[0] * g[2];
    cn1 = cn1 + cp1 * g[1] * g[2];
    ca1 = ca1 + cp1 * g[0] * g[2];
  }
  out[0] = cn * ca_ - ca * sa_;
  out[1] = cn * sa_ + ca * ca_;
  out[2] = cn1 * ca_ - ca1 * sa_;
  out[3] = cn1 * sa_ + ca1 * ca_;
}

// per wall face: x_f, cp (face-extrapolated pressure), cp (cell pressure),
// V_n with face-extrapolated velocity (src/io.f90:380-418); in batch
// contexts Mach number and free-stream pressure of the face's instance
__global__ void k_wall_cp(int nw, const int *__restrict__ wf, const int *__restrict__ bcell,
                          const double *__restrict__ geo, const double *__restrict__ W,
                          const double *__restrict__ grad, double mach, double pinf, const int *__restrict__ binst,
                          const double *__restrict__ mach_aoa, const double *__restrict__ pinf_inst,
                          double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nw) return;
  const int b = wf[k], c = bcell[b];
  double g[8], gr[8], w[4];
  ld8(geo, b, g);
  ld8(grad, c, gr);
  ld4(W, c, w);
  double M = mach, pi = pinf;
  if (binst) {
    const int q = binst[b];
    M = mach_aoa[2 * q];
    pi = pinf_inst[4 * q + 3];
  }
  const double pw = w[3] + g[3] * gr[3] + g[4] * gr[7];
  const double uw = w[1] + g[3] * gr[1] + g[4] * gr[5];
  const double vw = w[2] + g[3] * gr[2] + g[4] * gr[6];
  out[4 * k] = g[5];
  out[4 * k + 1] = 2.0 / (M * M) * (pw - pi);
  out[4 * k + 2] = 2.0 / (M * M) * (w[3] - pi);
  out[4 * k + 3] = uw * g[0] + vw * g[1];
}

__global__ void k_roe_batch(int n, double g, const double *__restrict__ pL, const double *__restrict__ pR,
                            const double *__restrict__ nxy, double *__restrict__ F, double *__restrict__ ws) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double a[4], b[4], f[4], w;
  ld4(pL, i, a);
  ld4(pR, i, b);
  roe_flux(g, a, b, nxy[2 * i], nxy[2 * i + 1], f, w);
  st4(F, i, f);
  ws[i] = w;
}

__global__ void k_fill(long n, double v, double *__restrict__ p) {
  const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
#include "fused.hpp"

}  // namespace

// Staging of the per-step outputs of fvs_time_integration_noalloc (4 residual
// norms and 15 vortex-error values per step, at most kNoAllocMaxSteps steps): a
// file-scope device array, so that entry point allocates nothing.  Every GPU
// holds its own copy at its own address until the process ends; host code
// uses the address hipGetSymbolAddress gives for the current device
// (scratch_address), never the array's name.
constexpr int kNoAllocMaxSteps = 4096;
constexpr size_t kScratchDoubles = size_t(kNoAllocMaxSteps) * (4 + kVerr);
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// device address of `scratch` on device dev, which must be the current device
int scratch_address(int dev, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) return fail(FVS_EHIP, "device index out of range");
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(FVS_EHIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

struct fvs_mesh {
  MeshPrep mp;
};

struct fvs_ctx {
  fvs_options opt{};
  fvs_mesh mesh;
  MeshPrep &mp = mesh.mp;
  Phys P{};
  int N = 0, Fi = 0, Fb = 0, GW = 0, NFW = 3;
  std::vector<int> perm, iperm;  // perm[new] = old
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  // device arrays
  double *U0 = nullptr, *Ub = nullptr, *Facc = nullptr, *W = nullptr, *grad = nullptr, *phi = nullptr,
         *dtl = nullptr, *FA = nullptr, *WSA = nullptr, *vol = nullptr, *cx = nullptr, *cy = nullptr;
  double *rout = nullptr, *wsout = nullptr, *part = nullptr, *small = nullptr;
  int2 *fi_cells = nullptr;
  double *fi_geo = nullptr;
  int *fb_cell = nullptr, *fb_type = nullptr;
  double *fb_geo = nullptr;
  int *gat = nullptr;
  int *nvrt = nullptr;
  double2 *loff = nullptr;
  int *soff = nullptr, *swid = nullptr, *snb = nullptr;
  double2 *scw = nullptr;
  double2 *ggcb0 = nullptr, *ggcbc = nullptr;
  int *ggcbnb = nullptr;
  // GGNB: node values F (NV nodes, ordered by their first device cell),
  // node incidences in slices, per cell gw node slots and coefedg
  int NV = 0, gwn = 0;
  int *nsoff = nullptr, *nswid = nullptr, *nscell = nullptr, *gcnode = nullptr;
  double *nscoef = nullptr, *gF = nullptr;
  double2 *gcedg = nullptr;
  long nnz_node = 0;
  int *cell_intr = nullptr, *wall = nullptr;
  int ncell_intr = 0, nwall = 0;
  // centroid-based LSQ gradient (k_grad_lsqc)
  bool lsqc = false;
  int pmode = 0;
  double2 *cxy = nullptr;
  double *ginv = nullptr;
  unsigned *sid = nullptr;
  ShiftTable shifts{};
  // fused flux + update (k_flux_update): per-block face lists, local gather
  bool fused = false;
  int maxf = 0;
  int *boff = nullptr, *bfaces = nullptr;
  short *lgat = nullptr;  // signed 1-based index into the block's face list
  // the same per-block lists for blocks of kSmallBlock cells (single meshes),
  // used by the fused flux kernel when small_blocks is set (autotuned)
  bool have_small = false, small_blocks = false;
  int maxf_s = 0;
  long nlisted_s = 0;
  int *boff_s = nullptr;
  short *lgat_s = nullptr;
  int2 *lrec_s = nullptr;
  double *lgeo_s = nullptr;
  long nlisted = 0;
  // the listed faces' records in list order: cells (interior: c1, c2;
  // boundary: cell, -1 - (type | instance << 2)) and geometry, 7 planes
  int2 *lrec = nullptr;
  double *lgeo = nullptr;
  double *Walt = nullptr;  // second primitive-state buffer
  // SSP-RK3 on the fused path with the staged-halo (or no) gradient: the
  // stages read conservative states (U0 -> Ub -> Uc -> U0) and no primitive
  // state is written; W is refreshed once after the steps
  double *Uc = nullptr;
  // XCD-aware block numbering per kernel class (fvs_autotune picks)
  bool xcd_grad = false, xcd_flux = false, xcd_upd = false, xcd_fused = true;
  // limiter values frozen at their last evaluation (steady convergence aid)
  bool limiter_frozen = false;
  // whole-stage kernel (k_stage): per-block C2 lists, local stencils and face cells
  bool stage = false, halo = false;  // k_stage in use; per-block data built (k_grad_lsqh)
  bool use_halo = true;              // gradient from the staged halo (autotuned)
  int max1 = 0, max2 = 0, max2n = 0;
  int *s2off = nullptr, *s2n1 = nullptr, *s2list = nullptr, *sl0 = nullptr, *slw = nullptr;
  long *sleo = nullptr;
  unsigned short *slid = nullptr;
  unsigned *blf = nullptr;
  long nslid = 0, nlist2 = 0, nslid_own = 0;
  // staged-halo gradient (k_grad_lsqh): per block, the cells outside it that
  // its stencils read (ascending), and every stencil entry as a 16-bit local
  // index (own cells 0..255, halo 256..) << 4 | periodic-shift code, two
  // entries per word: sliced ELL with the global stencil's widths, entries
  // 2kk, 2kk+1 of lane l of slice s at lcode[poff[s] + 64 kk + l]
  int *ghoff = nullptr, *ghlist = nullptr, *poff = nullptr;
  int *ghpad = nullptr, *ghpad_s = nullptr;  // the halo lists padded to maxh per block (-1 after the end)
  unsigned *lcode = nullptr;
  int maxh = 0;  // largest halo of a block
  long nghl = 0, nlcode = 0;
  // the same lists for blocks of kSmallBlock cells (single meshes), used by
  // the LSQ gradient when small_grad is set (autotuned)
  int *ghoff_s = nullptr, *ghlist_s = nullptr;
  unsigned *lcode_s = nullptr;
  int maxh_s = 0;
  bool have_small_grad = false, small_grad = false;
  // batch mode: per-instance frozen flags (dfrz) while only some are frozen
  int *dfrz = nullptr;
  bool frz_some = false;
  // batched instances (fvs_create_batch): contiguous device ranges
  int ninst = 0;
  std::vector<int> inst_of_real;     // reference order
  std::vector<int> inst_off_dev;     // device order, ninst+1
  std::vector<double> mach_aoa;      // 2 per instance
  // blist: nblk(N) entries of launch-block list, then 8 words per block of
  // active-cell bits (masked batches)
  int *dinst_off = nullptr, *dbinst = nullptr, *dwoff = nullptr, *blist = nullptr;
  double *dpinf_inst = nullptr, *d2buf = nullptr, *dmach_aoa = nullptr, *dforces = nullptr;
  int nact = 0;                      // active blocks (0 = all, no list)
  // step counter for norm outputs (graph replays), graph mode switch
  int *dctr = nullptr;
  bool use_graph = false;
  // device outputs of fvs_time_integration (per-step norms, vortex errors,
  // instance norms), kept across calls and grown on demand, so a call
  // allocates nothing once its step count has been seen
  double *dn_buf = nullptr, *dv_buf = nullptr, *di_buf = nullptr;
  size_t dn_cap = 0, dv_cap = 0, di_cap = 0;
  // the instantiated graph of the last graph-mode call and what its capture
  // baked in; launch_gen changes whenever a call changes what a step launches
  // (time step, limiter freezing, active blocks, autotuned launch choices)
  struct GraphKey {
    unsigned gen = 0;
    int gsteps = 0;
    const void *dn = nullptr, *di = nullptr, *w = nullptr, *walt = nullptr;
    bool operator==(const GraphKey &o) const {
      return gen == o.gen && gsteps == o.gsteps && dn == o.dn && di == o.di && w == o.w && walt == o.walt;
    }
  } gkey;
  hipGraphExec_t gexec = nullptr;
  unsigned launch_gen = 1;
  // persistent whole-step launches (fvs_set_persistent): grid barrier state
  // (count, generation) and its failure flag; compute units of the device
  bool use_persistent = false;
  unsigned *pbar = nullptr;
  int *perr = nullptr;
  int ncu = 0;
  bool persist_ran = false;
  size_t lds_max = 64 * 1024;  // LDS per workgroup the device allows
  // W does not hold the primitive state of U0 (the conservative-stage path
  // writes none); refreshed by the first caller that reads W
  bool w_stale = false;
  int nparts = 0;
  bool have_state = false;
  double last_ms = 0.0;
  // the device the context was created on, and the address of `scratch` there
  int device = 0;
  double *scratch_dev = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // per-launch event pairs while profiling (kernel class, start, stop)
  bool profiling = false;
  // phase timers (fvs_set_phase_timers): device ms per kernel class, summed
  // over the time-integration calls made while they are on
  bool phase_timing = false;
  double phase_ms[6] = {0, 0, 0, 0, 0, 0};
  std::vector<hipEvent_t> evpool;
  std::vector<int> evclass;
  size_t evused = 0;

  template <class T>
  int alloc(T *&p, size_t n) {
    void *q = nullptr;
    HIPC(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    allocs.push_back(q);
    p = static_cast<T *>(q);
    return 0;
  }
  template <class T>
  int upload(T *&p, const std::vector<T> &h) {
    if (int rc = alloc(p, h.size())) return rc;
    if (!h.empty()) HIPC(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    return 0;
  }
  // device buffer of at least n doubles, reallocated (doubling) only when n
  // exceeds what it holds
  int grow(double *&p, size_t &cap, size_t n) {
    if (n <= cap && p) return 0;
    const size_t want = std::max(n, 2 * cap);
    if (p) {
      HIPC(hipStreamSynchronize(stream));
      (void)hipFree(p);
      p = nullptr;
      cap = 0;
    }
    HIPC(hipMalloc(&p, std::max<size_t>(want, 1) * sizeof(double)));
    cap = want;
    return 0;
  }
  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gkey = GraphKey{};
  }
  ~fvs_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graph();
    for (double *q : {dn_buf, dv_buf, di_buf})
      if (q) (void)hipFree(q);
    for (void *q : allocs) (void)hipFree(q);
    for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Morton order of cell centroids: compact blocks of consecutive cells
std::vector<int> morton_order(const MeshPrep &m, const std::vector<int> &inst) {
  const int N = m.ncells;
  double xa = *std::min_element(m.cx.begin(), m.cx.end()), xb = *std::max_element(m.cx.begin(), m.cx.end());
  double ya = *std::min_element(m.cy.begin(), m.cy.end()), yb = *std::max_element(m.cy.begin(), m.cy.end());
  double s = std::max(xb - xa, yb - ya);
  if (!(s > 0)) s = 1.0;
  auto spread = [](uint64_t v) {
    v &= 0xffffffffull;
    v = (v | (v << 16)) & 0x0000ffff0000ffffull;
    v = (v | (v << 8)) & 0x00ff00ff00ff00ffull;
    v = (v | (v << 4)) & 0x0f0f0f0f0f0f0f0full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
  };
  // instances stay contiguous (instance-major), Morton order inside each
  std::vector<std::tuple<int, uint64_t, int>> key(N);
  for (int c = 0; c < N; ++c) {
    uint64_t ix = uint64_t((m.cx[c] - xa) / s * 4294967295.0), iy = uint64_t((m.cy[c] - ya) / s * 4294967295.0);
    key[c] = {inst.empty() ? 0 : inst[c], spread(ix) | (spread(iy) << 1), c};
  }
  std::sort(key.begin(), key.end());
  std::vector<int> perm(N);
  for (int c = 0; c < N; ++c) perm[c] = std::get<2>(key[c]);
  return perm;
}

// worker threads for host-side preparation: the cores this process may use,
// at most 16
int host_threads() {
  const unsigned n = std::thread::hardware_concurrency();
  return int(std::max(1u, std::min(16u, n)));
}

// [0, n) in contiguous parts of at least grain items, one per worker thread:
// parallel_parts gives their number, parallel_for runs fn(part, lo, hi) on
// each (fn must touch disjoint data per part)
long parallel_parts(long n, long grain) { return std::max(1L, std::min<long>(host_threads(), n / grain)); }
template <class F>
void parallel_for(long n, long grain, F fn) {
  const long nt = parallel_parts(n, grain);
  if (nt == 1) {
    fn(0L, 0L, n);
    return;
  }
  run_threads(int(nt), [&](int t) { fn(long(t), n * t / nt, n * (t + 1) / nt); });
}

// Recursive coordinate bisection of cell centroids, instance-major: a range
// [lo, hi) of device positions is cut at a multiple of 64 near its middle (at
// the largest power-of-two multiple of 64 within its middle third) along the
// longer side of its centroids' bounding box, down to 64-cell leaves. Cuts
// are placed by global position, so every aligned 64-, 128- and 256-cell block
// is one or a few compact pieces of one instance also where instances start
// off alignment; on graded meshes (O-grids) the pieces stay compact where a
// Z-curve over the bounding box does not. Ties and leaves are ordered by cell
// id, so the order does not depend on the selection algorithm or threads.
std::vector<int> rcb_order(const MeshPrep &m, const std::vector<int> &inst) {
  const long N = m.ncells;
  constexpr long kLeaf = 64;
  struct Pt {
    double x, y;
    int id;
  };
  auto key = [](double v) { return std::isnan(v) ? HUGE_VAL : v; };
  std::vector<Pt> p(N);
  for (long c = 0; c < N; ++c) p[c] = {key(m.cx[c]), key(m.cy[c]), int(c)};
  // instances are contiguous in the reference order and stay so
  std::vector<long> start{0};
  for (long c = 1; c < N; ++c)
    if (!inst.empty() && inst[c] != inst[c - 1]) start.push_back(c);
  start.push_back(N);
  const long ni = long(start.size()) - 1;
  auto cut = [](long lo, long hi) -> long {
    const long first = (lo / kLeaf + 1) * kLeaf;
    if (first >= hi) return -1;
    const long two_mid = lo + hi, n = hi - lo;
    long best = -1;
    for (long P = kLeaf;; P *= 2) {
      const long c = (two_mid / 2 / P) * P;
      long pick = -1;
      for (long k : {c, c + P})
        if (k > lo && k < hi && 3 * std::labs(2 * k - two_mid) <= n &&
            (pick < 0 || std::labs(2 * k - two_mid) < std::labs(2 * pick - two_mid)))
          pick = k;
      if (pick < 0) break;
      best = pick;
    }
    if (best < 0) {
      const long c = ((two_mid / 2 + kLeaf / 2) / kLeaf) * kLeaf;
      best = std::min(std::max(c, first), ((hi - 1) / kLeaf) * kLeaf);
    }
    return best;
  };
  auto by_id = [](const Pt &a, const Pt &b) { return a.id < b.id; };
  std::function<void(long, long, int)> run = [&](long lo, long hi, int spawn) {
    const long mid = cut(lo, hi);
    if (mid < 0) {
      std::sort(p.begin() + lo, p.begin() + hi, by_id);
      return;
    }
    double xa = HUGE_VAL, xb = -HUGE_VAL, ya = HUGE_VAL, yb = -HUGE_VAL;
    for (long i = lo; i < hi; ++i) {
      xa = std::min(xa, p[i].x), xb = std::max(xb, p[i].x);
      ya = std::min(ya, p[i].y), yb = std::max(yb, p[i].y);
    }
    const bool byx = xb - xa >= yb - ya;
    std::nth_element(p.begin() + lo, p.begin() + mid, p.begin() + hi, [byx](const Pt &a, const Pt &b) {
      const double ka = byx ? a.x : a.y, kb = byx ? b.x : b.y;
      return ka < kb || (ka == kb && a.id < b.id);
    });
    if (spawn > 0 && hi - lo > 65536) {
      run_threads(2, [&](int t) { t ? run(lo, mid, spawn - 1) : run(mid, hi, spawn - 1); });
    } else {
      run(lo, mid, 0);
      run(mid, hi, 0);
    }
  };
  const int nt = host_threads();
  if (ni >= nt) {
    parallel_for(ni, 1, [&](long, long k0, long k1) {
      for (long k = k0; k < k1; ++k) run(start[k], start[k + 1], 0);
    });
  } else {
    int spawn = 0;
    while ((2 << spawn) <= nt) ++spawn;
    for (long k = 0; k < ni; ++k) run(start[k], start[k + 1], spawn);
  }
  std::vector<int> perm(N);
  for (long i = 0; i < N; ++i) perm[i] = p[i].id;
  return perm;
}

// device cell order for the reorder option (perm[device position] = cell id)
std::vector<int> cell_order(const MeshPrep &m, int reorder, const std::vector<int> &inst) {
  if (reorder == 1) return morton_order(m, inst);
  if (reorder == 2) return rcb_order(m, inst);
  std::vector<int> perm(m.ncells);
  std::iota(perm.begin(), perm.end(), 0);
  return perm;
}

// sliced-ELL packing of a CSR stencil (cells in device order); false when
// the padded entry count does not fit the 32-bit slice offsets
bool pack_sell(int N, const std::vector<int> &perm, const std::vector<int> &iperm, const std::vector<int> &ptr,
               const std::vector<int> &col, const double *val2, std::vector<int> &off, std::vector<int> &wid,
               std::vector<int> &nb, std::vector<double2> &cw) {
  const int ns = (N + kSlice - 1) / kSlice;
  off.assign(ns + 1, 0);
  wid.assign(ns, 0);
  long total = 0;
  for (int s = 0; s < ns; ++s) {
    int w = 0;
    for (int l = 0; l < kSlice; ++l) {
      int i = s * kSlice + l;
      if (i < N) w = std::max(w, ptr[perm[i] + 1] - ptr[perm[i]]);
    }
    wid[s] = w;
    total += long(w) * kSlice;
    if (total > long(INT32_MAX)) return false;
    off[s + 1] = int(total);
  }
  nb.assign(off[ns], 0);
  cw.assign(val2 ? off[ns] : 0, make_double2(0.0, 0.0));
  parallel_for(ns, 64, [&](long, long s0, long s1) {
    for (long s = s0; s < s1; ++s)
      for (int l = 0; l < kSlice; ++l) {
        int i = int(s) * kSlice + l;
        for (int k = 0; k < wid[s]; ++k) {
          long e = off[s] + long(k) * kSlice + l;
          if (i >= N) { nb[e] = 0; continue; }
          int o = perm[i], b = ptr[o], n = ptr[o + 1] - b;
          if (k < n) {
            nb[e] = iperm[col[b + k]];
            if (val2) cw[e] = make_double2(val2[2 * size_t(b + k)], val2[2 * size_t(b + k) + 1]);
          } else {
            nb[e] = i;  // padding: the cell itself with zero weight
          }
        }
      }
  });
  return true;
}

// Per-block data of the whole-stage kernel (k_stage).  For block b: C1 =
// its own cells (in order) then the other cells of its listed faces
// (ascending); C2 = C1 then the remaining cells of the C1 stencils
// (ascending); the C1 stencils re-expressed as local C2 indices (<< 4, with
// the periodic-shift code in the low bits), in slices of 64 C1 cells padded
// with the cell itself; each listed face's cells as local indices.  The
// stencil entries keep the order of the global stencil, so the gradients are
// bit-identical to k_grad_lsqc's.  C.halo (k_grad_lsqh, the default
// gradient) is set when every block fits 4095 local cells and its staged
// halo 64 KiB; C.stage (k_stage) with kernel_path 2 when its whole LDS
// layout fits 64 KiB too.
int setup_stage(fvs_ctx &C, const std::vector<int> &boff, const std::vector<int> &bfaces,
                const std::vector<int2> &fic, const std::vector<int> &fbc, const std::vector<int> &off,
                const std::vector<int> &wid, const std::vector<int> &nb, const std::vector<unsigned> &code) {
  const int N = C.N, Fi = C.Fi, nbk = nblk(N);
  std::vector<int> s2off(nbk + 1, 0), s2n1(nbk, 0), s2list, sl0(nbk + 1, 0), slw;
  std::vector<long> sleo;
  std::vector<unsigned short> slid;
  std::vector<unsigned> blf(bfaces.size(), 0u);
  std::vector<int> lmap(N, -1), c1, extra;
  int max1 = 0, max2 = 0;
  long eo = 0, eo_own = 0;
  for (int b = 0; b < nbk; ++b) {
    const int i0 = b * kBlock, i1 = std::min(N, i0 + kBlock);
    c1.clear();
    extra.clear();
    for (int i = i0; i < i1; ++i) c1.push_back(i);
    for (int f = boff[b]; f < boff[b + 1]; ++f) {
      const int sl = bfaces[f];
      if (sl >= Fi) continue;
      const int a = fic[sl].x, c = fic[sl].y;
      if (a < i0 || a >= i1) extra.push_back(a);
      if (c < i0 || c >= i1) extra.push_back(c);
    }
    std::sort(extra.begin(), extra.end());
    extra.erase(std::unique(extra.begin(), extra.end()), extra.end());
    c1.insert(c1.end(), extra.begin(), extra.end());
    const int n1 = int(c1.size());
    for (int q = 0; q < n1; ++q) lmap[c1[q]] = q;
    extra.clear();
    for (int q = 0; q < n1; ++q) {
      const int i = c1[q], sg = i / kSlice, l = i % kSlice;
      for (int k = 0; k < wid[sg]; ++k) {
        const int j = nb[off[sg] + long(k) * kSlice + l];
        if (lmap[j] == -1) {
          lmap[j] = -2;
          extra.push_back(j);
        }
      }
    }
    std::sort(extra.begin(), extra.end());
    for (size_t k = 0; k < extra.size(); ++k) lmap[extra[k]] = n1 + int(k);
    const int n2 = n1 + int(extra.size());
    bool fits = n2 <= 4095;
    const int ns = (n1 + kSlice - 1) / kSlice;
    for (int tq = 0; tq < ns && fits; ++tq) {
      int w = 0;
      for (int q = tq * kSlice; q < std::min(n1, (tq + 1) * kSlice); ++q) w = std::max(w, wid[c1[q] / kSlice]);
      slw.push_back(w);
      sleo.push_back(eo);
      if (tq * kSlice < i1 - i0) eo_own += long(w) * kSlice;
      slid.resize(size_t(eo + long(w) * kSlice), 0);
      for (int l = 0; l < kSlice; ++l) {
        const int q = tq * kSlice + l;
        if (q >= n1) continue;
        const int i = c1[q], sg = i / kSlice, li = i % kSlice;
        for (int k = 0; k < w; ++k) {
          unsigned v = unsigned(q) << 4;  // padding: the cell itself, no shift
          if (k < wid[sg]) {
            const unsigned cd = code[off[sg] + long(k) * kSlice + li];
            v = (unsigned(lmap[int(cd & 0x07ffffffu)]) << 4) | (cd >> 27);
          }
          slid[size_t(eo + long(k) * kSlice + l)] = static_cast<unsigned short>(v);
        }
      }
      eo += long(w) * kSlice;
    }
    if (!fits) return 0;
    sl0[b + 1] = sl0[b] + ns;
    for (int f = boff[b]; f < boff[b + 1]; ++f) {
      const int sl = bfaces[f];
      if (sl < Fi) blf[f] = unsigned(lmap[fic[sl].x]) | (unsigned(lmap[fic[sl].y]) << 16);
      else blf[f] = unsigned(lmap[fbc[sl - Fi]]) | (0xffffu << 16);
    }
    // every local index the kernel reads stays inside the block's arrays
    for (int tq = 0; tq < ns; ++tq) {
      const int si = sl0[b] + tq;
      for (int l = 0; l < kSlice && tq * kSlice + l < n1; ++l)
        for (int k = 0; k < slw[si]; ++k)
          if (int(slid[size_t(sleo[si] + long(k) * kSlice + l)] >> 4) >= n2)
            return fail(FVS_EINVAL, "stage set-up: stencil entry outside the block's cell list");
    }
    for (int f = boff[b]; f < boff[b + 1]; ++f) {
      const int a = int(blf[f] & 0xffffu), c = int(blf[f] >> 16);
      if (a >= n1 || (bfaces[f] < Fi ? c >= n1 : c != 0xffff))
        return fail(FVS_EINVAL, "stage set-up: face cell outside the block's gradient cells");
    }
    for (int q = 0; q < i1 - i0; ++q)
      if (c1[q] != i0 + q) return fail(FVS_EINVAL, "stage set-up: own cells not first");
    s2list.insert(s2list.end(), c1.begin(), c1.end());
    s2list.insert(s2list.end(), extra.begin(), extra.end());
    s2off[b + 1] = int(s2list.size());
    s2n1[b] = n1;
    max1 = std::max(max1, n1);
    max2 = std::max(max2, n2);
    for (int c : c1) lmap[c] = -1;
    for (int c : extra) lmap[c] = -1;
  }
  max1 += max1 & 1;  // keep the LDS arrays after lP 16-byte aligned
  const int max2n = max2;
  if (size_t(6) * max2n * sizeof(double) > C.lds_max) return 0;
  // k_stage's block reduction reuses the staged-state area: 4 x 512 doubles
  max2 = std::max(max2, kStageThreads);
  const size_t lds = (size_t(6) * max2 + size_t(9) * max1 + size_t(5) * C.maxf) * sizeof(double);
  if (int rc = C.upload(C.s2off, s2off)) return rc;
  if (int rc = C.upload(C.s2n1, s2n1)) return rc;
  if (int rc = C.upload(C.s2list, s2list)) return rc;
  if (int rc = C.upload(C.sl0, sl0)) return rc;
  if (int rc = C.upload(C.slw, slw)) return rc;
  if (int rc = C.upload(C.sleo, sleo)) return rc;
  if (int rc = C.upload(C.slid, slid)) return rc;
  if (int rc = C.upload(C.blf, blf)) return rc;
  C.max1 = max1;
  C.max2 = max2;
  C.max2n = max2n;
  C.nslid = long(slid.size());
  C.nlist2 = long(s2list.size());
  C.nslid_own = eo_own;
  // the device's per-workgroup LDS limit (160 KiB on gfx950): blocks of
  // stretched cells (O-grids) have larger halos
  C.stage = lds <= C.lds_max;
  return 0;
}

// Lists of the staged-halo gradient k_grad_lsqh (see the fvs_ctx fields).
// Local entries keep the order of the global stencil, so the gradients are
// bit-identical to k_grad_lsqc's.  Enabled (C.halo) when every block's local
// cells number at most 4095 and their states and centres fit 64 KiB of LDS;
// built again for blocks of kSmallBlock cells on single meshes of at least
// 64 such blocks (C.have_small_grad).
int setup_grad_halo(fvs_ctx &C, const std::vector<int> &off, const std::vector<int> &wid,
                    const std::vector<unsigned> &code) {
  const int N = C.N;
  std::vector<int> poff(wid.size() + 1, 0);
  for (size_t sg = 0; sg < wid.size(); ++sg) poff[sg + 1] = poff[sg] + (wid[sg] + 1) / 2 * kSlice;  // <= off
  // the lists for blocks of BS cells; ok stays false when they do not fit
  auto build = [&](int BS, std::vector<int> &ghoff, std::vector<int> &ghlist, std::vector<unsigned> &lcode,
                   int &maxh, bool &ok) -> int {
    const int nbk = nblk(N, BS);
    ghoff.assign(nbk + 1, 0);
    ghlist.clear();
    lcode.assign(size_t(poff.back()), 0u);
    maxh = 0;
    ok = false;
    // blocks in contiguous parts on the host's cores (a block's cells are
    // whole 64-cell slices, so parts write disjoint code words); each part's
    // halo lists in block order, joined afterwards.  status: 1 = a block has
    // too many local cells, 2 = a stencil entry outside the local cells
    const long np = parallel_parts(nbk, 64);
    std::vector<std::vector<int>> ph(np);
    std::vector<int> pmax(np, 0), status(np, 0);
    parallel_for(nbk, 64, [&](long t, long b0, long b1) {
      std::vector<int> halo;
      for (long b = b0; b < b1 && !status[t]; ++b) {
        const int i0 = int(b) * BS, i1 = std::min(N, i0 + BS);
        halo.clear();
        for (int i = i0; i < i1; ++i) {
          const int sg = i / kSlice, l = i % kSlice;
          for (int k = 0; k < wid[sg]; ++k) {
            const int j = int(code[off[sg] + long(k) * kSlice + l] & 0x07ffffffu);
            if (j < i0 || j >= i1) halo.push_back(j);
          }
        }
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        if (BS + halo.size() > 4095) {
          status[t] = 1;
          break;
        }
        for (int i = i0; i < i1 && !status[t]; ++i) {
          const int sg = i / kSlice, l = i % kSlice;
          for (int k = 0; k < wid[sg]; ++k) {
            const long e = off[sg] + long(k) * kSlice + l;
            const int j = int(code[e] & 0x07ffffffu);
            int loc = j - i0;
            if (j < i0 || j >= i1) loc = BS + int(std::lower_bound(halo.begin(), halo.end(), j) - halo.begin());
            if (loc < 0 || loc >= BS + int(halo.size()) || (loc < BS && i0 + loc >= i1)) {
              status[t] = 2;
              break;
            }
            lcode[size_t(poff[sg]) + size_t(k / 2) * kSlice + l] |= ((unsigned(loc) << 4) | (code[e] >> 27))
                                                                    << (16 * (k & 1));
          }
        }
        ph[t].insert(ph[t].end(), halo.begin(), halo.end());
        ghoff[b + 1] = int(halo.size());
        pmax[t] = std::max(pmax[t], int(halo.size()));
      }
    });
    const int worst = *std::max_element(status.begin(), status.end());
    if (worst == 2) return fail(FVS_EINVAL, "halo set-up: stencil entry outside the block's local cells");
    if (worst == 1) return 0;
    for (int b = 0; b < nbk; ++b) ghoff[b + 1] += ghoff[b];
    ghlist.reserve(size_t(ghoff[nbk]));
    for (auto &v : ph) ghlist.insert(ghlist.end(), v.begin(), v.end());
    maxh = *std::max_element(pmax.begin(), pmax.end());
    ok = size_t(6) * (BS + maxh) * sizeof(double) <= 64 * 1024;
    if (ghlist.empty()) ghlist.push_back(0);
    return 0;
  };
  // every block's list at a fixed stride of maxh entries (-1 after its end),
  // so the staged-halo gradient finds its list without loading its offset
  auto pad = [&](const std::vector<int> &ghoff, const std::vector<int> &ghlist, int maxh, int *&dst) -> int {
    const long nbk = long(ghoff.size()) - 1;
    std::vector<int> p(size_t(nbk) * size_t(std::max(maxh, 1)), -1);
    parallel_for(nbk, 256, [&](long, long b0, long b1) {
      for (long b = b0; b < b1; ++b)
        std::copy(ghlist.begin() + ghoff[b], ghlist.begin() + ghoff[b + 1], p.begin() + b * long(maxh));
    });
    return C.upload(dst, p);
  };
  std::vector<int> ghoff, ghlist;
  std::vector<unsigned> lcode;
  int maxh = 0;
  bool ok = false;
  if (int rc = build(kBlock, ghoff, ghlist, lcode, maxh, ok)) return rc;
  if (!ok) return 0;
  if (int rc = C.upload(C.ghoff, ghoff)) return rc;
  if (int rc = C.upload(C.ghlist, ghlist)) return rc;
  if (int rc = pad(ghoff, ghlist, maxh, C.ghpad)) return rc;
  if (int rc = C.upload(C.poff, poff)) return rc;
  if (int rc = C.upload(C.lcode, lcode)) return rc;
  C.maxh = maxh;
  C.nghl = ghoff.back();
  C.nlcode = poff.back();
  C.halo = true;
  if (C.ninst == 0 && N >= 64 * kSmallBlock) {
    if (int rc = build(kSmallBlock, ghoff, ghlist, lcode, maxh, ok)) return rc;
    if (ok) {
      if (int rc = C.upload(C.ghoff_s, ghoff)) return rc;
      if (int rc = C.upload(C.ghlist_s, ghlist)) return rc;
      if (int rc = pad(ghoff, ghlist, maxh, C.ghpad_s)) return rc;
      if (int rc = C.upload(C.lcode_s, lcode)) return rc;
      C.maxh_s = maxh;
      C.have_small_grad = true;
    }
  }
  return 0;
}

int setup_device(fvs_ctx &C) {
  MeshPrep &m = C.mp;
  const int N = C.N = m.ncells;
  C.perm.resize(N);
  C.perm = cell_order(m, C.opt.reorder, C.inst_of_real);
  C.iperm.resize(N);
  for (int i = 0; i < N; ++i) C.iperm[C.perm[i]] = i;
  const auto &P = C.perm, &IP = C.iperm;
  C.NFW = (m.nquad > 0) ? 4 : 3;
  // cells
  std::vector<double> vol(N), cx(N), cy(N);
  std::vector<int> nv(N);
  for (int i = 0; i < N; ++i) vol[i] = m.vol[P[i]], cx[i] = m.cx[P[i]], cy[i] = m.cy[P[i]], nv[i] = m.nvrt[P[i]];
  if (int rc = C.upload(C.vol, vol)) return rc;
  if (int rc = C.upload(C.cx, cx)) return rc;
  if (int rc = C.upload(C.cy, cy)) return rc;
  if (int rc = C.upload(C.nvrt, nv)) return rc;
  // interior faces in ascending reference edge id
  const int Fi = C.Fi = int(m.edge_intr.size());
  std::vector<int2> fic(Fi);
  std::vector<double> fig(8 * size_t(Fi), 0.0);
  std::vector<int> slot_of_edge(m.nedges, -1);
  parallel_for(Fi, 4096, [&](long, long f0, long f1) {
    for (long f = f0; f < f1; ++f) {
      const int e = m.edge_intr[f], c1 = m.ec1[e], c2 = m.ec2[e];
      slot_of_edge[e] = int(f);
      fic[f] = make_int2(IP[c1], IP[c2]);
      const double xf = m.ex[e], yf = m.ey[e];
      const double xr = m.cx[c2] + m.esx[e], yr = m.cy[c2] + m.esy[e];
      double *g = &fig[8 * size_t(f)];
      g[0] = m.enx[e]; g[1] = m.eny[e]; g[2] = m.ea[e];
      g[3] = xf - m.cx[c1]; g[4] = yf - m.cy[c1];
      g[5] = xf - xr; g[6] = yf - yr;
    }
  });
  if (int rc = C.upload(C.fi_cells, fic)) return rc;
  if (int rc = C.upload(C.fi_geo, fig)) return rc;
  // boundary faces: per boundary, edge i paired with listed cell i
  const int Fb = C.Fb = int(m.bnd_edges.size());
  std::vector<int> fbc(Fb), fbt(Fb), wall;
  std::vector<double> fbg(8 * size_t(Fb), 0.0);
  // each cell's signed face slots (reference cell order, CSR): its interior
  // faces in ascending slot, then its boundary faces in ascending slot
  std::vector<int> cf_ptr(size_t(N) + 1, 0), cf_fill;
  for (int f = 0; f < Fi; ++f) {
    const int e = m.edge_intr[f];
    ++cf_ptr[m.ec1[e] + 1];
    ++cf_ptr[m.ec2[e] + 1];
  }
  for (int ib = 0; ib < m.nb; ++ib)
    for (int k = 0; k < m.bnd_nedges[ib]; ++k) ++cf_ptr[m.bnd_cells[m.bnd_cptr[ib] + k] + 1];
  for (int c = 0; c < N; ++c) cf_ptr[c + 1] += cf_ptr[c];
  std::vector<int> cell_faces(static_cast<size_t>(cf_ptr[N]));
  cf_fill.assign(cf_ptr.begin(), cf_ptr.end() - 1);
  for (int f = 0; f < Fi; ++f) {
    const int e = m.edge_intr[f];
    cell_faces[cf_fill[m.ec1[e]]++] = f + 1;
    cell_faces[cf_fill[m.ec2[e]]++] = -(f + 1);
  }
  int b = 0;
  for (int ib = 0; ib < m.nb; ++ib)
    for (int k = 0; k < m.bnd_nedges[ib]; ++k, ++b) {
      const int e = m.bnd_edges[m.bnd_eptr[ib] + k];
      const int c = m.bnd_cells[m.bnd_cptr[ib] + k];
      fbc[b] = IP[c];
      fbt[b] = m.bnd_type[ib];
      double *g = &fbg[8 * size_t(b)];
      g[0] = m.enx[e]; g[1] = m.eny[e]; g[2] = m.ea[e];
      g[3] = m.ex[e] - m.cx[c]; g[4] = m.ey[e] - m.cy[c];
      g[5] = m.ex[e]; g[6] = m.ey[e];
      cell_faces[cf_fill[c]++] = Fi + b + 1;
      if (m.bnd_type[ib] == FVS_BC_SLIP_WALL) wall.push_back(b);
    }
  if (int rc = C.upload(C.fb_cell, fbc)) return rc;
  if (int rc = C.upload(C.fb_type, fbt)) return rc;
  if (int rc = C.upload(C.fb_geo, fbg)) return rc;
  C.nwall = int(wall.size());
  if (int rc = C.upload(C.wall, wall)) return rc;
  // gather table: interior faces are already in ascending edge id because
  // slots follow edge_intr; boundary slots follow them
  int gw = 1;
  for (int c = 0; c < N; ++c) gw = std::max(gw, cf_ptr[c + 1] - cf_ptr[c]);
  C.GW = gw;
  std::vector<int> gat(size_t(gw) * N, 0);
  parallel_for(N, 4096, [&](long, long i0, long i1) {
    for (long i = i0; i < i1; ++i) {
      const int c = P[i], a = cf_ptr[c], n = cf_ptr[c + 1] - a;
      for (int k = 0; k < n; ++k) gat[size_t(k) * N + i] = cell_faces[size_t(a) + k];
    }
  });
  if (int rc = C.upload(C.gat, gat)) return rc;
  // per-block face lists for the fused flux + update kernel: every face slot
  // touching block b's cells, ascending (interior slots, then boundary slots),
  // and each cell's faces as signed local indices in the gather order
  std::vector<int> h_boff, h_bfaces;
  // per-block lists for blocks of bs cells: boff, bfaces, lgat and the
  // largest list (maxf)
  auto block_lists = [&](int bs, std::vector<int> &boff, std::vector<int> &bfaces, std::vector<short> &lgat,
                         int &maxf) {
    const int nbk = nblk(N, bs);
    boff.assign(nbk + 1, 0);
    bfaces.clear();
    lgat.assign(size_t(gw) * N, 0);
    // blocks in contiguous parts on the host's cores; each part's lists in
    // block order, joined afterwards
    const long np = parallel_parts(nbk, 64);
    std::vector<std::vector<int>> pf(np);
    std::vector<int> pmax(np, 0);
    parallel_for(nbk, 64, [&](long t, long b0, long b1) {
      std::vector<int> loc;
      for (long b = b0; b < b1; ++b) {
        loc.clear();
        const int i0 = int(b) * bs, i1 = std::min(N, i0 + bs);
        for (int i = i0; i < i1; ++i)
          for (int k = 0; k < gw; ++k) {
            const int sg = gat[size_t(k) * N + i];
            if (sg) loc.push_back((sg > 0 ? sg : -sg) - 1);
          }
        std::sort(loc.begin(), loc.end());
        loc.erase(std::unique(loc.begin(), loc.end()), loc.end());
        for (int i = i0; i < i1; ++i)
          for (int k = 0; k < gw; ++k) {
            const int sg = gat[size_t(k) * N + i];
            if (!sg) continue;
            const int slot = (sg > 0 ? sg : -sg) - 1;
            const int li = int(std::lower_bound(loc.begin(), loc.end(), slot) - loc.begin()) + 1;
            lgat[size_t(k) * N + i] = short(sg > 0 ? li : -li);
          }
        pf[t].insert(pf[t].end(), loc.begin(), loc.end());
        boff[b + 1] = int(loc.size());
        pmax[t] = std::max(pmax[t], int(loc.size()));
      }
    });
    for (int b = 0; b < nbk; ++b) boff[b + 1] += boff[b];
    bfaces.reserve(size_t(boff[nbk]));
    for (auto &v : pf) bfaces.insert(bfaces.end(), v.begin(), v.end());
    maxf = *std::max_element(pmax.begin(), pmax.end());
  };
  // the listed faces' records in list order: cells and 7 geometry planes
  auto upload_records = [&](const std::vector<int> &bfaces, int2 *&dlrec, double *&dlgeo) -> int {
    const size_t L = bfaces.size();
    std::vector<int2> lrec(L);
    std::vector<double> lgeo(7 * L);
    parallel_for(long(L), 4096, [&](long, long f0, long f1) {
      for (long f = f0; f < f1; ++f) {
        const int sl = bfaces[f];
        const double *g;
        if (sl < Fi) {
          lrec[f] = fic[sl];
          g = &fig[8 * size_t(sl)];
        } else {
          const int bq = sl - Fi;
          const int inst = C.ninst > 0 ? C.inst_of_real[P[fbc[bq]]] : 0;
          lrec[f] = make_int2(fbc[bq], -1 - (fbt[bq] | (inst << 2)));
          g = &fbg[8 * size_t(bq)];
        }
        for (int k = 0; k < 7; ++k) lgeo[size_t(k) * L + f] = g[k];
      }
    });
    if (int rc = C.upload(dlrec, lrec)) return rc;
    return C.upload(dlgeo, lgeo);
  };
  // (the fused update gathers at most 4 faces per cell: triangles and quads)
  if ((C.opt.kernel_path == 0 || C.opt.kernel_path == 2 || C.opt.kernel_path == 3) && gw <= 4) {
    std::vector<int> boff, bfaces;
    std::vector<short> lgat;
    int maxf = 0;
    block_lists(kBlock, boff, bfaces, lgat, maxf);
    // LDS of k_flux_update: face fluxes (at least the 4 x 256 reduction) +
    // the block's gradients and states
    if (size_t(std::max(5 * maxf, 4 * kBlock)) * sizeof(double) + 12 * kBlock * sizeof(double) <= 64 * 1024) {
      C.maxf = maxf;
      C.nlisted = long(bfaces.size());
      if (int rc = C.upload(C.boff, boff)) return rc;
      if (int rc = C.upload(C.bfaces, bfaces)) return rc;
      if (int rc = C.upload(C.lgat, lgat)) return rc;
      if (int rc = upload_records(bfaces, C.lrec, C.lgeo)) return rc;
      C.fused = true;
      h_boff.swap(boff);
      h_bfaces.swap(bfaces);
    }
    // the same lists for blocks of kSmallBlock cells, a launch shape the
    // autotuner tries (single meshes of at least 64 blocks: batches keep the
    // 256-cell blocks their instance masks are built for)
    if (C.fused && C.ninst == 0 && C.opt.kernel_path != 2 && N >= 64 * kSmallBlock) {
      block_lists(kSmallBlock, boff, bfaces, lgat, maxf);
      C.maxf_s = maxf;
      C.nlisted_s = long(bfaces.size());
      if (int rc = C.upload(C.boff_s, boff)) return rc;
      if (int rc = C.upload(C.lgat_s, lgat)) return rc;
      if (int rc = upload_records(bfaces, C.lrec_s, C.lgeo_s)) return rc;
      C.have_small = true;
    }
  }
  // limiter face offsets (xf - xc) per local edge, in the cell's own frame
  if (C.opt.limiter > 0) {
    std::vector<double2> lo(4 * size_t(N), make_double2(0.0, 0.0));
    parallel_for(N, 4096, [&](long, long i0, long i1) {
      for (long i = i0; i < i1; ++i) {
        const int c = P[i];
        for (int k = 0; k < m.nvrt[c]; ++k) {
          const int e = m.cedge[4 * c + k];
          double xc, yc;
          side_centroid(m, e, c, xc, yc);
          lo[size_t(k) * N + i] = make_double2(m.ex[e] - xc, m.ey[e] - yc);
        }
      }
    });
    if (int rc = C.upload(C.loff, lo)) return rc;
  }
  // LSQ stencil (gradient and/or limiter)
  if (!m.lsq_ptr.empty()) {
    std::vector<double> cw(2 * m.lsq_cell.size());
    parallel_for(long(m.lsq_cell.size()), 4096, [&](long, long k0, long k1) {
      for (long k = k0; k < k1; ++k) {
        cw[2 * k] = m.lsq_coef[2 * k] * m.lsq_w[k];
        cw[2 * k + 1] = m.lsq_coef[2 * k + 1] * m.lsq_w[k];
      }
    });
    std::vector<int> off, wid, nb;
    std::vector<double2> c2;
    if (!pack_sell(N, P, IP, m.lsq_ptr, m.lsq_cell, cw.data(), off, wid, nb, c2))
      return fail(FVS_EINVAL, "mesh too large: LSQ stencil entries exceed 32-bit offsets");
    if (int rc = C.upload(C.soff, off)) return rc;
    if (int rc = C.upload(C.swid, wid)) return rc;
    if (int rc = C.upload(C.snb, nb)) return rc;
    if (int rc = C.upload(C.scw, c2)) return rc;
    // centroid form: neighbour id | shift code << 27, per-cell G^-1
    if (C.opt.grad_method == 3 && C.opt.kernel_path != 1 && N < (1 << 27)) {
      std::vector<double> dsx, dsy;  // distinct periodic shifts
      std::vector<unsigned> code(nb.size(), 0u);
      bool ok = true;
      for (int s = 0; s < int(wid.size()) && ok; ++s)
        for (int k = 0; k < wid[s] && ok; ++k)
          for (int l = 0; l < kSlice && ok; ++l) {
            const long e = off[s] + long(k) * kSlice + l;
            const int i = s * kSlice + l;
            unsigned sh = 0;
            if (i < N && m.periodic) {
              const int ci = P[i], cj = P[nb[e]];
              double sx = 0.0, sy = 0.0;
              for (int a = 0; a < m.nvrt[ci] && sx == 0.0 && sy == 0.0; ++a)
                for (int b = 0; b < m.nvrt[cj]; ++b)
                  if (m.cnodec[4 * ci + a] == m.cnodec[4 * cj + b]) {
                    sx = m.nxy[2 * m.cnode[4 * ci + a]] - m.nxy[2 * m.cnode[4 * cj + b]];
                    sy = m.nxy[2 * m.cnode[4 * ci + a] + 1] - m.nxy[2 * m.cnode[4 * cj + b] + 1];
                    break;
                  }
              if (sx != 0.0 || sy != 0.0) {
                size_t q = 0;
                while (q < dsx.size() && !(dsx[q] == sx && dsy[q] == sy)) ++q;
                if (q == dsx.size()) {
                  if (q == 15) { ok = false; break; }
                  dsx.push_back(sx);
                  dsy.push_back(sy);
                }
                sh = unsigned(q + 1);
              }
            }
            code[e] = unsigned(nb[e]) | (sh << 27);
          }
      if (ok) {
        for (size_t q = 0; q < dsx.size(); ++q) C.shifts.sx[q] = dsx[q], C.shifts.sy[q] = dsy[q];
        std::vector<double2> cxy(N);
        std::vector<double> gi(3 * size_t(N), 0.0);  // planes: G^-1 11, 12, 22
        const double p = C.opt.lsq_p;
        parallel_for(N, 4096, [&](long, long i0, long i1) {
          for (long i = i0; i < i1; ++i) {
            const int c = P[i];
            cxy[i] = make_double2(m.cx[c], m.cy[c]);
            // G = D^T D with rows w*(dx,dy), w = 1/d^p (src/gradient_lsq.f90:159-189)
            double g11 = 0, g12 = 0, g22 = 0;
            for (int k = m.lsq_ptr[c]; k < m.lsq_ptr[c + 1]; ++k) {
              const int j = m.lsq_cell[k];
              const double w = m.lsq_w[k];
              double sx = 0.0, sy = 0.0;
              if (m.periodic)
                for (int a = 0; a < m.nvrt[c] && sx == 0.0 && sy == 0.0; ++a)
                  for (int b = 0; b < m.nvrt[j]; ++b)
                    if (m.cnodec[4 * c + a] == m.cnodec[4 * j + b]) {
                      sx = m.nxy[2 * m.cnode[4 * c + a]] - m.nxy[2 * m.cnode[4 * j + b]];
                      sy = m.nxy[2 * m.cnode[4 * c + a] + 1] - m.nxy[2 * m.cnode[4 * j + b] + 1];
                      break;
                    }
              const double d1 = w * ((m.cx[j] + sx) - m.cx[c]), d2 = w * ((m.cy[j] + sy) - m.cy[c]);
              g11 = g11 + d1 * d1;
              g12 = g12 + d1 * d2;
              g22 = g22 + d2 * d2;
            }
            const double det = g11 * g22 - g12 * g12;
            gi[size_t(i)] = 1.0 / det * g22;
            gi[size_t(N) + i] = -(1.0 / det * g12);
            gi[2 * size_t(N) + i] = 1.0 / det * g11;
          }
        });
        C.pmode = (p == 0.0) ? 0 : (p == 1.0 ? 1 : (p == 2.0 ? 2 : 3));
        if (int rc = C.upload(C.cxy, cxy)) return rc;
        if (int rc = C.upload(C.ginv, gi)) return rc;
        if (int rc = C.upload(C.sid, code)) return rc;
        C.lsqc = true;
        if (C.opt.recon != 1 && C.opt.kernel_path != 3)
          if (int rc = setup_grad_halo(C, off, wid, code)) return rc;
        if (C.fused && C.opt.recon != 1 && C.opt.kernel_path == 2)
          if (int rc = setup_stage(C, h_boff, h_bfaces, fic, fbc, off, wid, nb, code)) return rc;
      }
    }
    // a limited Green-Gauss gradient takes its bounds from a staged halo of
    // the same (node-neighbour) stencil: codes are the device ids (the
    // bounds need no periodic shift)
    if ((C.opt.grad_method == 1 || C.opt.grad_method == 2) && C.opt.limiter > 0 && C.opt.recon != 1 &&
        C.opt.kernel_path != 3 && N < (1 << 27)) {
      std::vector<unsigned> code(nb.begin(), nb.end());
      if (int rc = setup_grad_halo(C, off, wid, code)) return rc;
    }
  }
  if (C.opt.grad_method == 1) {
    std::vector<double2> c0(N), cn(size_t(C.NFW) * N, make_double2(0.0, 0.0));
    std::vector<int> nb(size_t(C.NFW) * N);
    for (int i = 0; i < N; ++i) {
      const int c = P[i];
      c0[i] = make_double2(m.ggcb_coef0[2 * c], m.ggcb_coef0[2 * c + 1]);
      for (int k = 0; k < C.NFW; ++k) {
        if (k < m.nvrt[c]) {
          nb[size_t(k) * N + i] = IP[m.ggcb_ptr[4 * c + k]];
          cn[size_t(k) * N + i] = make_double2(m.ggcb_coefnb[8 * c + 2 * k], m.ggcb_coefnb[8 * c + 2 * k + 1]);
        } else {
          nb[size_t(k) * N + i] = i;
        }
      }
    }
    if (int rc = C.upload(C.ggcb0, c0)) return rc;
    if (int rc = C.upload(C.ggcbnb, nb)) return rc;
    if (int rc = C.upload(C.ggcbc, cn)) return rc;
  }
  if (C.opt.grad_method == 2) {
    // nodes that touch a cell, ordered by their first device cell so that a
    // slice of nodes reads a compact range of cells and vice versa
    const int Vn = m.nnodes;
    std::vector<int> first(Vn, INT32_MAX);
    for (int v = 0; v < Vn; ++v)
      for (int q = m.n2c_ptr[v]; q < m.n2c_ptr[v + 1]; ++q) first[v] = std::min(first[v], IP[m.n2c[q]]);
    std::vector<int> nodes;
    for (int v = 0; v < Vn; ++v)
      if (first[v] < INT32_MAX) nodes.push_back(v);
    std::stable_sort(nodes.begin(), nodes.end(), [&](int a, int b) { return first[a] < first[b]; });
    const int NV = int(nodes.size());
    std::vector<int> dnode(Vn, -1);
    for (int k = 0; k < NV; ++k) dnode[nodes[k]] = k;
    // node incidences in 64-node slices, the reference's cell order per node
    const int ns = (NV + kSlice - 1) / kSlice;
    std::vector<int> noff(ns + 1, 0), nwid(ns, 0);
    long total = 0;
    for (int sl = 0; sl < ns; ++sl) {
      int w = 0;
      for (int l = 0; l < kSlice && sl * kSlice + l < NV; ++l) {
        const int v = nodes[sl * kSlice + l];
        w = std::max(w, m.n2c_ptr[v + 1] - m.n2c_ptr[v]);
      }
      nwid[sl] = w;
      total += long(w) * kSlice;
      if (total > long(INT32_MAX)) return fail(FVS_EINVAL, "mesh too large: GGNB node incidences exceed 32-bit offsets");
      noff[sl + 1] = int(total);
    }
    std::vector<int> ncell(size_t(total), 0);
    std::vector<double> ncoef(size_t(total), 0.0);
    for (int sl = 0; sl < ns; ++sl)
      for (int l = 0; l < kSlice; ++l) {
        const int k = sl * kSlice + l;
        const int v = k < NV ? nodes[k] : nodes[0];
        const int b = m.n2c_ptr[v], n = m.n2c_ptr[v + 1] - b;
        for (int e = 0; e < nwid[sl]; ++e) {
          const size_t at = size_t(noff[sl]) + size_t(e) * kSlice + l;
          ncell[at] = IP[m.n2c[b + (e < n ? e : 0)]];  // padding: a cell of the node, weight 0
          ncoef[at] = (k < NV && e < n) ? m.ggnb_ncoef[b + e] : 0.0;
        }
      }
    // per cell: its nodes (the reference's local order) and coefedg
    int gwn = 0;
    for (int c = 0; c < N; ++c) gwn = std::max(gwn, m.nvrt[c]);
    std::vector<int> cnode(size_t(gwn) * N, 0);
    std::vector<double2> cedg(size_t(gwn) * N, make_double2(0.0, 0.0));
    for (int i = 0; i < N; ++i) {
      const int c = P[i];
      for (int k = 0; k < gwn; ++k) {
        const bool real = k < m.nvrt[c];
        cnode[size_t(k) * N + i] = dnode[m.cnodec[4 * c + (real ? k : 0)]];
        if (real) cedg[size_t(k) * N + i] = make_double2(m.ggnb_edg[8 * size_t(c) + 2 * k], m.ggnb_edg[8 * size_t(c) + 2 * k + 1]);
      }
    }
    C.NV = NV;
    C.gwn = gwn;
    C.nnz_node = long(m.n2c.size());
    if (int rc = C.upload(C.nsoff, noff)) return rc;
    if (int rc = C.upload(C.nswid, nwid)) return rc;
    if (int rc = C.upload(C.nscell, ncell)) return rc;
    if (int rc = C.upload(C.nscoef, ncoef)) return rc;
    if (int rc = C.upload(C.gcnode, cnode)) return rc;
    if (int rc = C.upload(C.gcedg, cedg)) return rc;
    if (int rc = C.alloc(C.gF, 4 * size_t(NV))) return rc;
  }
  // interior cells (vortex error), device ids, reference order
  std::vector<int> ci(m.cell_intr.size());
  for (size_t k = 0; k < ci.size(); ++k) ci[k] = IP[m.cell_intr[k]];
  C.ncell_intr = int(ci.size());
  if (int rc = C.upload(C.cell_intr, ci)) return rc;
  // state and work arrays
  const size_t n4 = 4 * size_t(N);
  if (int rc = C.alloc(C.U0, n4)) return rc;
  if (int rc = C.alloc(C.Ub, n4)) return rc;
  if (int rc = C.alloc(C.Facc, n4)) return rc;
  if (int rc = C.alloc(C.W, n4)) return rc;
  if (int rc = C.alloc(C.Walt, n4)) return rc;
  if (int rc = C.alloc(C.Uc, n4)) return rc;
  if (int rc = C.alloc(C.grad, 2 * n4)) return rc;
  if (int rc = C.alloc(C.phi, N)) return rc;
  if (int rc = C.alloc(C.dtl, N)) return rc;
  if (int rc = C.alloc(C.FA, 4 * size_t(Fi + Fb))) return rc;
  if (int rc = C.alloc(C.WSA, size_t(Fi + Fb))) return rc;
  if (int rc = C.alloc(C.rout, n4)) return rc;
  if (int rc = C.alloc(C.wsout, N)) return rc;
  C.nparts = std::max(nblk(N, kSmallBlock), nblk(std::max(C.ncell_intr, 1)));
  if (int rc = C.alloc(C.part, 13 * size_t(C.nparts))) return rc;
  if (int rc = C.alloc(C.small, 64)) return rc;
  if (int rc = C.alloc(C.dctr, 1)) return rc;
  if (int rc = C.alloc(C.pbar, 2)) return rc;
  if (int rc = C.alloc(C.perr, 1)) return rc;
  if (C.ninst > 0) {
    // per-instance ranges in device order, free streams, wall-face ranges
    std::vector<int> off(C.ninst + 1, 0), binst(Fb), woff(C.ninst + 1, 0);
    for (int i = 0; i < N; ++i) off[C.inst_of_real[P[i]] + 1]++;
    for (int k = 0; k < C.ninst; ++k) off[k + 1] += off[k];
    for (int i = 1; i < N; ++i)
      if (C.inst_of_real[P[i]] < C.inst_of_real[P[i - 1]]) return fail(FVS_EINVAL, "instances not contiguous");
    C.inst_off_dev = off;
    for (int q = 0; q < Fb; ++q) binst[q] = C.inst_of_real[P[fbc[q]]];
    std::vector<double> pinf(4 * size_t(C.ninst));
    const double kPi = 3.14159265358979323846;
    for (int k = 0; k < C.ninst; ++k) {
      pinf[4 * k] = 1.0;
      pinf[4 * k + 1] = C.mach_aoa[2 * k] * std::cos(C.mach_aoa[2 * k + 1] * (kPi / 180.0));
      pinf[4 * k + 2] = C.mach_aoa[2 * k] * std::sin(C.mach_aoa[2 * k + 1] * (kPi / 180.0));
      pinf[4 * k + 3] = 1.0 / C.opt.gamma;
    }
    std::vector<int> wsorted(wall);
    std::stable_sort(wsorted.begin(), wsorted.end(), [&](int a, int b2) { return binst[a] < binst[b2]; });
    for (int w : wsorted) woff[binst[w] + 1]++;
    for (int k = 0; k < C.ninst; ++k) woff[k + 1] += woff[k];
    if (int rc = C.upload(C.dinst_off, off)) return rc;
    if (int rc = C.upload(C.dbinst, binst)) return rc;
    if (int rc = C.upload(C.dpinf_inst, pinf)) return rc;
    if (int rc = C.upload(C.dmach_aoa, C.mach_aoa)) return rc;
    if (int rc = C.upload(C.dwoff, woff)) return rc;
    HIPC(hipMemcpyAsync(C.wall, wsorted.data(), wsorted.size() * sizeof(int), hipMemcpyHostToDevice, C.stream));
    if (int rc = C.alloc(C.d2buf, n4)) return rc;
    if (int rc = C.alloc(C.dforces, 4 * size_t(C.ninst))) return rc;
    if (int rc = C.alloc(C.blist, 9 * size_t(nblk(N)))) return rc;
  }
  // phi = 1 (no limiter) or 0 (first order); grad = 0 for first order
  const double phi0 = (C.opt.recon == 1) ? 0.0 : 1.0;
  hipLaunchKernelGGL(k_fill, dim3(nblk(N)), dim3(kBlock), 0, C.stream, long(N), phi0, C.phi);
  hipLaunchKernelGGL(k_fill, dim3(nblk(8L * N)), dim3(kBlock), 0, C.stream, 8L * N, 0.0, C.grad);
  hipLaunchKernelGGL(k_fill, dim3(nblk(N)), dim3(kBlock), 0, C.stream, long(N), C.opt.dt, C.dtl);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(C.stream));
  return 0;
}

// kernel classes timed by fvs_kernel_times
enum { KC_GRAD = 0, KC_FLUX_INT = 1, KC_FLUX_BND = 2, KC_UPDATE = 3, KC_NORM = 4, KC_FUSED = 5, KC_N = 6 };

struct ProfScope {
  fvs_ctx &C;
  int cls;
  bool on;
  ProfScope(fvs_ctx &c, int k) : C(c), cls(k), on(c.profiling) {
    if (!on) return;
    if (C.evused + 2 > C.evpool.size()) {
      for (int q = 0; q < 64; ++q) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) { on = false; return; }
        C.evpool.push_back(e);
      }
    }
    (void)hipEventRecord(C.evpool[C.evused], C.stream);
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(C.evpool[C.evused + 1], C.stream);
    C.evclass.push_back(cls);
    C.evused += 2;
  }
};

// gradient (and limiter) of the primitive state W, or with cons of the
// conservative state src (staged-halo kernel only)
int launch_gradient(fvs_ctx &C, bool with_limiter, const double *src = nullptr) {
  const int N = C.N;
  const bool lim = with_limiter && C.opt.limiter > 0 && C.opt.recon != 1 && !C.limiter_frozen;
  const int lt = C.opt.limiter;
  dim3 g(xcd_grid(nblk(N))), b(kBlock);
  const bool gg = C.opt.grad_method == 1 || C.opt.grad_method == 2;
  if (src && !(C.halo && C.use_halo) && !(gg && !lim))
    return fail(FVS_ESTATE, "conservative gradient input needs the staged halo or an unlimited Green-Gauss gradient");
  const double *Wg = src ? src : C.W;
  const int cons = src ? 1 : 0;
  ProfScope ps(C, KC_GRAD);
  const bool ghalo = gg && lim && C.halo && C.use_halo;  // limited Green-Gauss, bounds from the staged halo
  const size_t glds = size_t(4) * (kBlock + C.maxh) * sizeof(double);
  if (C.opt.grad_method == 3 && C.halo && C.use_halo) {
    // 128-cell blocks (single meshes only: batches keep the 256-cell blocks
    // their instance masks are built for)
    const bool small = C.small_grad && C.have_small_grad && !C.nact;
    if (C.nact) g = dim3(C.nact);
    const int nb = small ? nblk(N, kSmallBlock) : nblk(N);
    if (small) g = dim3(xcd_grid(nb)), b = dim3(kSmallBlock);
    const int nl = small ? kSmallBlock + C.maxh_s : kBlock + C.maxh;
    const size_t lds = size_t(6) * nl * sizeof(double);
    const int *ghpad = small ? C.ghpad_s : C.ghpad;
    const unsigned *lcode = small ? C.lcode_s : C.lcode;
#define FVS_LSQH(L, PM, BS)                                                                                      \
  hipLaunchKernelGGL((k_grad_lsqh<L, PM, BS>), g, b, lds, C.stream, N, xcd_arg(nb, C.xcd_grad), nl,             \
                     src ? src : C.W, src ? 1 : 0, C.opt.gamma, ghpad, C.poff, C.swid, lcode, C.cxy, C.ginv,        \
                     C.shifts, C.opt.lsq_p, C.grad, C.phi, lt, C.nvrt, C.loff, C.vol, C.nact ? C.blist : nullptr,   \
                     C.frz_some ? C.dinst_off : nullptr, C.dfrz, C.ninst)
#define FVS_LSQH_PM(L, BS)                     \
  do {                                         \
    if (C.pmode == 0) FVS_LSQH(L, 0, BS);      \
    else if (C.pmode == 1) FVS_LSQH(L, 1, BS); \
    else if (C.pmode == 2) FVS_LSQH(L, 2, BS); \
    else FVS_LSQH(L, 3, BS);                   \
  } while (0)
    if (small) {
      if (lim) FVS_LSQH_PM(1, kSmallBlock);
      else FVS_LSQH_PM(0, kSmallBlock);
    } else {
      if (lim) FVS_LSQH_PM(1, kBlock);
      else FVS_LSQH_PM(0, kBlock);
    }
#undef FVS_LSQH_PM
#undef FVS_LSQH
  } else if (C.lsqc) {
    if (C.nact) g = dim3(C.nact);
    const int nb = nblk(N);
#define FVS_LSQC(L, PM)                                                                                          \
  hipLaunchKernelGGL((k_grad_lsqc<L, PM>), g, b, 0, C.stream, N, xcd_arg(nb, C.xcd_grad), C.W, C.cxy, C.ginv, C.soff, C.swid, C.sid,  \
                     C.shifts, C.opt.lsq_p, C.grad, C.phi, lt, C.nvrt, C.loff, C.vol, C.nact ? C.blist : nullptr,      \
                     C.frz_some ? C.dinst_off : nullptr, C.dfrz, C.ninst)
    if (lim) {
      if (C.pmode == 0) FVS_LSQC(1, 0);
      else if (C.pmode == 1) FVS_LSQC(1, 1);
      else if (C.pmode == 2) FVS_LSQC(1, 2);
      else FVS_LSQC(1, 3);
    } else {
      if (C.pmode == 0) FVS_LSQC(0, 0);
      else if (C.pmode == 1) FVS_LSQC(0, 1);
      else if (C.pmode == 2) FVS_LSQC(0, 2);
      else FVS_LSQC(0, 3);
    }
#undef FVS_LSQC
  } else if (C.opt.grad_method == 3) {
    if (lim)
      hipLaunchKernelGGL(k_grad_lsq<1>, g, b, 0, C.stream, xcd_arg(nblk(N), C.xcd_grad), N, C.W, C.soff, C.swid, C.snb, C.scw, C.grad, C.phi, lt,
                         C.nvrt, C.loff, C.vol);
    else
      hipLaunchKernelGGL(k_grad_lsq<0>, g, b, 0, C.stream, xcd_arg(nblk(N), C.xcd_grad), N, C.W, C.soff, C.swid, C.snb, C.scw, C.grad, C.phi, lt,
                         C.nvrt, C.loff, C.vol);
  } else if (C.opt.grad_method == 1) {
    if (ghalo)
      hipLaunchKernelGGL(k_grad_ggcb_h, g, b, glds, C.stream, xcd_arg(nblk(N), C.xcd_grad), N, C.NFW, Wg, cons,
                         C.opt.gamma, C.ggcb0,
                         C.ggcbnb, C.ggcbc, C.vol, C.grad, C.phi, lt, C.nvrt, C.loff, C.ghoff, C.ghlist, C.poff,
                         C.swid, C.lcode);
    else if (lim)
      hipLaunchKernelGGL(k_grad_ggcb<1>, g, b, 0, C.stream, xcd_arg(nblk(N), C.xcd_grad), N, C.NFW, C.W, 0, C.opt.gamma, C.ggcb0, C.ggcbnb, C.ggcbc, C.vol, C.grad,
                         C.phi, lt, C.nvrt, C.loff, C.soff, C.swid, C.snb);
    else
      hipLaunchKernelGGL(k_grad_ggcb<0>, g, b, 0, C.stream, xcd_arg(nblk(N), C.xcd_grad), N, C.NFW, Wg, cons, C.opt.gamma, C.ggcb0, C.ggcbnb, C.ggcbc, C.vol, C.grad,
                         C.phi, lt, C.nvrt, C.loff, C.soff, C.swid, C.snb);
  } else {
    hipLaunchKernelGGL(k_ggnb_nodes, dim3(nblk(C.NV)), b, 0, C.stream, C.NV, Wg, cons, C.opt.gamma,
                       C.nsoff, C.nswid, C.nscell, C.nscoef, C.gF);
    if (ghalo)
      hipLaunchKernelGGL(k_grad_ggnb_h, g, b, glds, C.stream, xcd_arg(nblk(N), C.xcd_grad), N, C.gwn, Wg, cons,
                         C.opt.gamma, C.gcnode, C.gcedg, C.gF, C.vol, C.grad, C.phi, lt, C.nvrt, C.loff, C.ghoff, C.ghlist,
                         C.poff, C.swid, C.lcode);
    else if (lim)
      hipLaunchKernelGGL(k_grad_ggnb<1>, g, b, 0, C.stream, xcd_arg(nblk(N), C.xcd_grad), N, C.gwn, C.W, C.gcnode,
                         C.gcedg, C.gF, C.vol, C.grad, C.phi, lt, C.nvrt, C.loff, C.soff, C.swid, C.snb);
    else
      hipLaunchKernelGGL(k_grad_ggnb<0>, g, b, 0, C.stream, xcd_arg(nblk(N), C.xcd_grad), N, C.gwn, C.W, C.gcnode,
                         C.gcedg, C.gF, C.vol, C.grad, C.phi, lt, C.nvrt, C.loff, C.soff, C.swid, C.snb);
  }
  HIPC(hipGetLastError());
  return 0;
}

// gradient + limiter + face fluxes of the current primitive state W
int launch_fluxes(fvs_ctx &C, double t) {
  if (C.opt.recon != 1)
    if (int rc = launch_gradient(C, true)) return rc;
  if (C.Fi > 0) {
    ProfScope ps(C, KC_FLUX_INT);
    hipLaunchKernelGGL(k_flux_int, dim3(xcd_grid(nblk(C.Fi))), dim3(kBlock), 0, C.stream,
                       xcd_arg(nblk(C.Fi), C.xcd_flux), C.Fi, C.P, C.fi_cells, C.fi_geo,
                       C.W, C.grad, C.phi, C.FA, C.WSA);
  }
  if (C.Fb > 0) {
    ProfScope ps(C, KC_FLUX_BND);
    hipLaunchKernelGGL(k_flux_bnd, dim3(nblk(C.Fb)), dim3(kBlock), 0, C.stream, C.Fb, C.Fi, C.P, t, C.fb_cell,
                       C.fb_type, C.fb_geo, C.W, C.grad, C.phi, C.FA, C.WSA, C.dbinst, C.dpinf_inst);
  }
  HIPC(hipGetLastError());
  return 0;
}

int launch_update(fvs_ctx &C, const Upd &A) {
  ProfScope ps(C, KC_UPDATE);
  hipLaunchKernelGGL(k_update, dim3(xcd_grid(nblk(C.N))), dim3(kBlock), 0, C.stream, xcd_arg(nblk(C.N), C.xcd_upd),
                     C.N, C.GW, A, C.gat, C.FA, C.WSA, C.vol,
                     C.U0, C.Ub, C.Facc, C.W, C.dtl, C.rout, C.wsout, C.part);
  HIPC(hipGetLastError());
  return 0;
}

// fused flux + gather + update of one stage; swaps the primitive buffers
// uin / uout (SSP-RK3, conservative stages): the stage's input state and the
// buffer its new stage state goes to; without them the stage reads W and
// writes the next primitive state to Walt (then swapped)
int launch_flux_update(fvs_ctx &C, const Upd &A, double t, const double *uin = nullptr, double *uout = nullptr) {
  const bool cons = uin != nullptr;
  if (C.opt.recon != 1)
    if (int rc = launch_gradient(C, true, uin)) return rc;
  {
    ProfScope ps(C, KC_FUSED);
    // the wave-speed array only where the stage sets the local time step
    const int per = (A.steady && A.first) ? 5 : 4;
    const int phi_one = (C.opt.limiter == 0 && C.opt.recon != 1) ? 1 : 0;
    if (C.small_blocks) {
      const int nb = nblk(C.N, kSmallBlock);
      const size_t lds = size_t(std::max(per * C.maxf_s, 4 * kSmallBlock)) * sizeof(double);
      hipLaunchKernelGGL(k_flux_update<kSmallBlock>, dim3(xcd_grid(nb)), dim3(kSmallBlock), lds, C.stream, C.N,
                         xcd_arg(nb, C.xcd_fused), C.GW, C.maxf_s, A, C.P, t, C.boff_s, C.lrec_s, C.lgeo_s,
                         C.nlisted_s, C.lgat_s, cons ? uin : C.W, C.grad, C.phi, C.vol, C.U0, cons ? uout : C.Ub,
                         C.Facc, cons ? nullptr : C.Walt, C.dtl, C.part, nullptr, nullptr,
                         A.last ? C.d2buf : nullptr, phi_one, cons ? 1 : 0);
    } else {
      const int nb = nblk(C.N);
      const size_t lds = size_t(std::max(per * C.maxf, 4 * kBlock)) * sizeof(double);
      hipLaunchKernelGGL(k_flux_update<kBlock>, dim3(C.nact ? C.nact : xcd_grid(nb)), dim3(kBlock), lds, C.stream,
                         C.N, xcd_arg(nb, C.xcd_fused), C.GW, C.maxf, A, C.P, t, C.boff, C.lrec, C.lgeo, C.nlisted,
                         C.lgat, cons ? uin : C.W, C.grad, C.phi, C.vol, C.U0, cons ? uout : C.Ub, C.Facc,
                         cons ? nullptr : C.Walt, C.dtl, C.part, C.nact ? C.blist : nullptr,
                         C.ninst > 0 ? C.dpinf_inst : nullptr, A.last ? C.d2buf : nullptr, phi_one, cons ? 1 : 0);
    }
  }
  HIPC(hipGetLastError());
  if (!cons) std::swap(C.W, C.Walt);
  return 0;
}

// arguments of the whole-stage block body for the current context state
StageArgs stage_args(const fvs_ctx &C, double t) {
  StageArgs a;
  a.N = C.N;
  a.nb = xcd_arg(nblk(C.N), C.xcd_fused);
  a.gw = C.GW;
  a.maxf = C.maxf;
  a.max1 = C.max1;
  a.max2 = C.max2;
  a.P = C.P;
  a.t = t;
  a.boff = C.boff;
  a.bfaces = C.bfaces;
  a.blf = C.blf;
  a.lgat = C.lgat;
  a.Fi = C.Fi;
  a.fgeo = C.fi_geo;
  a.btype = C.fb_type;
  a.bgeo = C.fb_geo;
  a.s2off = C.s2off;
  a.s2n1 = C.s2n1;
  a.s2list = C.s2list;
  a.sl0 = C.sl0;
  a.slw = C.slw;
  a.sleo = C.sleo;
  a.slid = C.slid;
  a.cxy = C.cxy;
  a.ginv = C.ginv;
  a.st = C.shifts;
  a.p = C.opt.lsq_p;
  a.phisrc = C.opt.limiter > 0 ? 1 : 0;
  a.phi = C.phi;
  a.lt = C.opt.limiter;
  a.nvrt = C.nvrt;
  a.loff = C.loff;
  a.vol = C.vol;
  a.ioff = C.frz_some ? C.dinst_off : nullptr;
  a.frz = C.dfrz;
  a.ninst = C.ninst;
  a.Ua = C.U0;
  a.Ub = C.Ub;
  a.Facc = C.Facc;
  a.dtl = C.dtl;
  a.part = C.part;
  a.blist = C.nact ? C.blist : nullptr;
  a.binst = C.dbinst;
  a.pinf_inst = C.dpinf_inst;
  return a;
}

size_t stage_lds(const fvs_ctx &C) {
  return (size_t(6) * C.max2 + size_t(9) * C.max1 + size_t(5) * C.maxf) * sizeof(double);
}

// whole stage in one kernel (k_stage); swaps the primitive buffers
int launch_stage(fvs_ctx &C, const Upd &A, double t) {
  ProfScope ps(C, KC_FUSED);
  const int nb = nblk(C.N);
  const size_t lds = stage_lds(C);
  const bool lim = C.opt.limiter > 0 && !C.limiter_frozen;
  const dim3 g(C.nact ? C.nact : xcd_grid(nb)), b(kStageThreads);
  const StageArgs a = stage_args(C, t);
  double *d2 = A.last ? C.d2buf : nullptr;
#define FVS_STAGE(L, PM) hipLaunchKernelGGL((k_stage<L, PM>), g, b, lds, C.stream, a, A, C.W, C.Walt, d2)
  if (lim) {
    if (C.pmode == 0) FVS_STAGE(1, 0);
    else if (C.pmode == 1) FVS_STAGE(1, 1);
    else if (C.pmode == 2) FVS_STAGE(1, 2);
    else FVS_STAGE(1, 3);
  } else {
    if (C.pmode == 0) FVS_STAGE(0, 0);
    else if (C.pmode == 1) FVS_STAGE(0, 1);
    else if (C.pmode == 2) FVS_STAGE(0, 2);
    else FVS_STAGE(0, 3);
  }
#undef FVS_STAGE
  HIPC(hipGetLastError());
  std::swap(C.W, C.Walt);
  return 0;
}

int launch_vortex_err(fvs_ctx &C, double t, double *dst) {
  const int n = C.ncell_intr;
  if (n == 0) return 0;
  const int nb = nblk(n);
  hipLaunchKernelGGL(k_vortex_err, dim3(nb), dim3(kBlock), 0, C.stream, n, C.P, t, C.cell_intr, C.cx, C.cy, C.U0,
                     C.part);
  hipLaunchKernelGGL(k_vortex_final, dim3(1), dim3(64), 0, C.stream, nb, n, C.part, C.cell_intr, C.cx, C.cy,
                     C.iperm[0], dst);
  HIPC(hipGetLastError());
  return 0;
}

// host <-> device cell-order permutation of per-cell records of width k
template <class T>
void to_device_order(const fvs_ctx &C, const T *src, std::vector<T> &dst, int k) {
  dst.resize(size_t(C.N) * k);
  for (int i = 0; i < C.N; ++i)
    for (int j = 0; j < k; ++j) dst[size_t(i) * k + j] = src[size_t(C.perm[i]) * k + j];
}
template <class T>
void to_ref_order(const fvs_ctx &C, const std::vector<T> &src, T *dst, int k) {
  for (int i = 0; i < C.N; ++i)
    for (int j = 0; j < k; ++j) dst[size_t(C.perm[i]) * k + j] = src[size_t(i) * k + j];
}

int need_state(const fvs_ctx *C) {
  if (!C) return fail(FVS_EINVAL, "null context");
  if (!C->have_state) return fail(FVS_ESTATE, "no state: call fvs_set_state or fvs_initialize first");
  return 0;
}

// W := primitive state of U0
void refresh_prim(fvs_ctx &C) {
  hipLaunchKernelGGL(k_prim, dim3(nblk(C.N)), dim3(kBlock), 0, C.stream, C.N, C.opt.gamma, C.U0, C.W);
  C.w_stale = false;
}
// the primitive state of U0 in W, where a conservative-stage run left it stale
void ensure_prim(fvs_ctx &C) {
  if (C.w_stale) refresh_prim(C);
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
// An exception thrown by host code inside an entry point (an allocation of
// the preparation's or a transfer's host vectors that fails, a worker thread
// that cannot start) ends at the C-ABI as an error code and message, never
// as a C++ exception crossing into the caller.
static int host_exception(const char *fn) {
  try {
    throw;
  } catch (const std::bad_alloc &) {
    return fail(FVS_ENOMEM, std::string(fn) + ": host memory exhausted");
  } catch (const std::exception &e) {
    return fail(FVS_ESTATE, std::string(fn) + ": " + e.what());
  } catch (...) {
    return fail(FVS_ESTATE, std::string(fn) + ": host exception");
  }
}

extern "C" {

const char *fvs_last_error(void) { return g_err.c_str(); }

}  // extern "C"

// error text for the other translation units of the library (mlp.hip)
namespace fvs {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace fvs

static int create_impl(fvs_ctx **out, int nnodes, const double *xy, int ntri, int nquad, const int *cell_nodes,
                       const int *node_canon, int nb, const int *bnd_ncells, const int *bnd_types,
                       const int *bnd_cells, const fvs_options *opt, int ninst, const int *cell_offsets,
                       const double *mach_aoa) {
  if (!out || !xy || !cell_nodes || !opt || (nb > 0 && (!bnd_ncells || !bnd_types)))
    return fail(FVS_EINVAL, "fvs_create: null argument");
  *out = nullptr;
  if (opt->recon < 1 || opt->recon > 3) return fail(FVS_EINVAL, "recon must be 1, 2 or 3");
  if (opt->limiter < 0 || opt->limiter > 3) return fail(FVS_EINVAL, "limiter must be 0..3");
  if (opt->rk_scheme < 0 || opt->rk_scheme > 3) return fail(FVS_EINVAL, "unknown rk_scheme");
  if (opt->reorder < 0 || opt->reorder > 2) return fail(FVS_EINVAL, "reorder must be 0, 1 or 2");
  if (opt->rk_scheme == FVS_RK_CLASSIC && (opt->rk_order < 1 || opt->rk_order > 4))
    return fail(FVS_EINVAL, "rk_order must be 1..4");
  if (!(opt->gamma > 1.0)) return fail(FVS_EINVAL, "gamma must exceed 1");
  if (ntri < 0 || nquad < 0 || ntri + nquad < 1 || nnodes < 3) return fail(FVS_EINVAL, "fvs_create: empty mesh");
  if (ninst > 0) {
    const int N = ntri + nquad;
    if (!cell_offsets || !mach_aoa || cell_offsets[0] != 0 || cell_offsets[ninst] != N)
      return fail(FVS_EINVAL, "fvs_create_batch: cell_offsets must run from 0 to ncells");
    for (int k = 0; k < ninst; ++k)
      if (cell_offsets[k + 1] <= cell_offsets[k])
        return fail(FVS_EINVAL, "fvs_create_batch: cell_offsets must increase (every instance has cells)");
  }
  int devs = 0;
  if (hipGetDeviceCount(&devs) != hipSuccess || devs < 1) return fail(FVS_EHIP, "no HIP device available");
  std::unique_ptr<fvs_ctx> C(new fvs_ctx());
  C->opt = *opt;
  if (ninst > 0) {
    const int N = ntri + nquad;
    C->ninst = ninst;
    C->inst_of_real.resize(N);
    for (int k = 0; k < ninst; ++k)
      for (int c = cell_offsets[k]; c < cell_offsets[k + 1]; ++c) C->inst_of_real[c] = k;
    C->mach_aoa.assign(mach_aoa, mach_aoa + 2 * size_t(ninst));
  }
  SchemeFlags fl;
  fl.grad_method = opt->grad_method;
  fl.lsq_nn = opt->lsq_nn;
  fl.lsq_p = opt->lsq_p;
  fl.limiter = opt->limiter;
  // a batch of cells of one kind is prepared instance by instance on the
  // host's cores
  std::string err = (ninst > 0 && (ntri == 0 || nquad == 0))
                        ? build_mesh_batch(C->mp, ninst, cell_offsets, nnodes, xy, ntri, nquad, cell_nodes, nb,
                                           bnd_ncells, bnd_types, bnd_cells, fl, host_threads())
                        : build_mesh(C->mp, nnodes, xy, ntri, nquad, cell_nodes, node_canon, nb, bnd_ncells,
                                     bnd_types, bnd_cells, fl);
  if (!err.empty()) {
    C.reset();
    return fail(FVS_EINVAL, err);
  }
  const double kPi = 3.14159265358979323846;
  Phys &P = C->P;
  P.gamma = opt->gamma;
  P.kappa = (opt->recon == 3) ? opt->kappa : 0.0;  // src/input.f90:258-271
  P.pinf[0] = 1.0;                                  // src/data_solution.f90:56-59
  P.pinf[1] = opt->mach * std::cos(opt->aoa_deg * (kPi / 180.0));
  P.pinf[2] = opt->mach * std::sin(opt->aoa_deg * (kPi / 180.0));
  P.pinf[3] = 1.0 / opt->gamma;
  P.vx0 = opt->vortex_x0;
  P.vy0 = opt->vortex_y0;
  P.vK = opt->vortex_K;
  for (int k = 0; k < 4; ++k) P.vinf[k] = opt->vortex_inf[k];
  P.lvortex = opt->lvortex;
  P.limiter = opt->limiter;
  if (hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking) != hipSuccess) {
    C.reset();
    return fail(FVS_EHIP, "hipStreamCreate failed");
  }
  (void)hipEventCreate(&C->ev0);
  (void)hipEventCreate(&C->ev1);
  int rc0 = hipGetDevice(&C->device) == hipSuccess ? 0 : fail(FVS_EHIP, "hipGetDevice failed");
  if (!rc0 && hipDeviceGetAttribute(&C->ncu, hipDeviceAttributeMultiprocessorCount, C->device) != hipSuccess)
    rc0 = fail(FVS_EHIP, "hipDeviceGetAttribute failed");
  int lmax = 0;
  if (!rc0 && hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, C->device) != hipSuccess)
    rc0 = fail(FVS_EHIP, "hipDeviceGetAttribute failed");
  if (!rc0 && lmax > 0) C->lds_max = size_t(lmax);
  if (!rc0) rc0 = scratch_address(C->device, &C->scratch_dev);
  if (rc0) {
    std::string e = g_err;
    C.reset();
    g_err = e;
    return rc0;
  }
  if (int rc = setup_device(*C)) {
    std::string e = g_err;
    C.reset();
    g_err = e;
    return rc;
  }
  *out = C.release();
  return 0;
}

extern "C" {

int fvs_create(fvs_ctx **out, int nnodes, const double *xy, int ntri, int nquad, const int *cell_nodes,
               const int *node_canon, int nb, const int *bnd_ncells, const int *bnd_types, const int *bnd_cells,
               const fvs_options *opt) try {
  return create_impl(out, nnodes, xy, ntri, nquad, cell_nodes, node_canon, nb, bnd_ncells, bnd_types, bnd_cells, opt,
                     0, nullptr, nullptr);
} catch (...) {
  return host_exception("fvs_create");
}

int fvs_create_batch(fvs_ctx **out, int ninst, const int *cell_offsets, const double *mach_aoa, int nnodes,
                     const double *xy, int ntri, int nquad, const int *cell_nodes, int nb, const int *bnd_ncells,
                     const int *bnd_types, const int *bnd_cells, const fvs_options *opt) try {
  if (ninst <= 0) return fail(FVS_EINVAL, "fvs_create_batch: ninst must be positive");
  return create_impl(out, nnodes, xy, ntri, nquad, cell_nodes, nullptr, nb, bnd_ncells, bnd_types, bnd_cells, opt,
                     ninst, cell_offsets, mach_aoa);
} catch (...) {
  return host_exception("fvs_create_batch");
}

int fvs_destroy(fvs_ctx *ctx) try {
  delete ctx;
  return 0;
} catch (...) {
  return host_exception("fvs_destroy");
}

int fvs_sizes(const fvs_ctx *C, int *o) try {
  if (!C || !o) return fail(FVS_EINVAL, "null argument");
  const MeshPrep &m = C->mp;
  o[0] = m.ncells; o[1] = m.nedges; o[2] = m.nnodes; o[3] = int(m.edge_intr.size());
  o[4] = m.nedges_bndr; o[5] = int(m.cell_intr.size()); o[6] = m.ncells_bndr; o[7] = m.nb;
  return 0;
} catch (...) {
  return host_exception("fvs_sizes");
}

int fvs_mesh_create(fvs_mesh **out, int nnodes, const double *xy, int ntri, int nquad, const int *cell_nodes,
                    const int *node_canon, int nb, const int *bnd_ncells, const int *bnd_types, const int *bnd_cells,
                    int grad_method, int lsq_nn, double lsq_p, int limiter) try {
  if (!out || !xy || !cell_nodes || (nb > 0 && (!bnd_ncells || !bnd_types)))
    return fail(FVS_EINVAL, "fvs_mesh_create: null argument");
  *out = nullptr;
  SchemeFlags fl;
  fl.grad_method = grad_method;
  fl.lsq_nn = lsq_nn;
  fl.lsq_p = lsq_p;
  fl.limiter = limiter;
  std::unique_ptr<fvs_mesh> M(new fvs_mesh());
  std::string err = build_mesh(M->mp, nnodes, xy, ntri, nquad, cell_nodes, node_canon, nb, bnd_ncells, bnd_types,
                               bnd_cells, fl);
  if (!err.empty()) return fail(FVS_EINVAL, err);
  *out = M.release();
  return 0;
} catch (...) {
  return host_exception("fvs_mesh_create");
}

int fvs_mesh_create_batch(fvs_mesh **out, int ninst, const int *cell_offsets, int nnodes, const double *xy,
                          int ntri, int nquad, const int *cell_nodes, int nb, const int *bnd_ncells,
                          const int *bnd_types, const int *bnd_cells, int grad_method, int lsq_nn, double lsq_p,
                          int limiter, int nthreads) try {
  if (!out || !xy || !cell_nodes || !cell_offsets || (nb > 0 && (!bnd_ncells || !bnd_types || !bnd_cells)))
    return fail(FVS_EINVAL, "fvs_mesh_create_batch: null argument");
  *out = nullptr;
  if (ninst < 1 || cell_offsets[0] != 0 || cell_offsets[ninst] != ntri + nquad)
    return fail(FVS_EINVAL, "fvs_mesh_create_batch: cell_offsets must run from 0 to ncells");
  for (int k = 0; k < ninst; ++k)
    if (cell_offsets[k + 1] <= cell_offsets[k])
      return fail(FVS_EINVAL, "fvs_mesh_create_batch: cell_offsets must increase (every instance has cells)");
  SchemeFlags fl;
  fl.grad_method = grad_method;
  fl.lsq_nn = lsq_nn;
  fl.lsq_p = lsq_p;
  fl.limiter = limiter;
  std::unique_ptr<fvs_mesh> M(new fvs_mesh());
  std::string err = build_mesh_batch(M->mp, ninst, cell_offsets, nnodes, xy, ntri, nquad, cell_nodes, nb, bnd_ncells,
                                     bnd_types, bnd_cells, fl, nthreads > 0 ? nthreads : host_threads());
  if (!err.empty()) return fail(FVS_EINVAL, err);
  *out = M.release();
  return 0;
} catch (...) {
  return host_exception("fvs_mesh_create_batch");
}

int fvs_mesh_destroy(fvs_mesh *M) try {
  delete M;
  return 0;
} catch (...) {
  return host_exception("fvs_mesh_destroy");
}

const fvs_mesh *fvs_ctx_mesh(const fvs_ctx *C) { return C ? &C->mesh : nullptr; }

int fvs_mesh_sizes(const fvs_mesh *M, int *o) try {
  if (!M || !o) return fail(FVS_EINVAL, "null argument");
  const MeshPrep &m = M->mp;
  o[0] = m.ncells; o[1] = m.nedges; o[2] = m.nnodes; o[3] = int(m.edge_intr.size());
  o[4] = m.nedges_bndr; o[5] = int(m.cell_intr.size()); o[6] = m.ncells_bndr; o[7] = m.nb;
  return 0;
} catch (...) {
  return host_exception("fvs_mesh_sizes");
}

int fvs_mesh_topology(const fvs_mesh *M, int *n1, int *n2, int *c1, int *c2, double *g7, double *xyv3, int *nghbre4,
                      int *edge4) try {
  if (!M) return fail(FVS_EINVAL, "null mesh");
  const MeshPrep &m = M->mp;
  const size_t E = m.nedges, N = m.ncells;
  if (n1) std::copy(m.en1.begin(), m.en1.end(), n1);
  if (n2) std::copy(m.en2.begin(), m.en2.end(), n2);
  if (c1) std::copy(m.ec1.begin(), m.ec1.end(), c1);
  if (c2) std::copy(m.ec2.begin(), m.ec2.end(), c2);
  if (g7) {
    const std::vector<double> *v[7] = {&m.ex, &m.ey, &m.ea, &m.enx, &m.eny, &m.esx, &m.esy};
    for (int k = 0; k < 7; ++k) std::copy(v[k]->begin(), v[k]->end(), g7 + k * E);
  }
  if (xyv3) {
    std::copy(m.cx.begin(), m.cx.end(), xyv3);
    std::copy(m.cy.begin(), m.cy.end(), xyv3 + N);
    std::copy(m.vol.begin(), m.vol.end(), xyv3 + 2 * N);
  }
  if (nghbre4) std::copy(m.nghbre.begin(), m.nghbre.end(), nghbre4);
  if (edge4) std::copy(m.cedge.begin(), m.cedge.end(), edge4);
  return 0;
} catch (...) {
  return host_exception("fvs_mesh_topology");
}

int fvs_mesh_lists(const fvs_mesh *M, int *edge_intr, int *cell_intr, int *bnd_nedges, int *bnd_edges) try {
  if (!M) return fail(FVS_EINVAL, "null mesh");
  const MeshPrep &m = M->mp;
  if (edge_intr) std::copy(m.edge_intr.begin(), m.edge_intr.end(), edge_intr);
  if (cell_intr) std::copy(m.cell_intr.begin(), m.cell_intr.end(), cell_intr);
  if (bnd_nedges) std::copy(m.bnd_nedges.begin(), m.bnd_nedges.end(), bnd_nedges);
  if (bnd_edges) std::copy(m.bnd_edges.begin(), m.bnd_edges.end(), bnd_edges);
  return 0;
} catch (...) {
  return host_exception("fvs_mesh_lists");
}

int fvs_mesh_cell_order(const fvs_mesh *M, int reorder, int ninst, const int *cell_offsets, int *perm) try {
  if (!M || !perm || (ninst > 0 && !cell_offsets)) return fail(FVS_EINVAL, "fvs_mesh_cell_order: null argument");
  if (reorder < 0 || reorder > 2) return fail(FVS_EINVAL, "reorder must be 0, 1 or 2");
  const MeshPrep &m = M->mp;
  std::vector<int> inst;
  if (ninst > 0) {
    if (cell_offsets[0] != 0 || cell_offsets[ninst] != m.ncells)
      return fail(FVS_EINVAL, "fvs_mesh_cell_order: cell_offsets must run from 0 to ncells");
    inst.resize(m.ncells);
    for (int k = 0; k < ninst; ++k) {
      if (cell_offsets[k + 1] <= cell_offsets[k])
        return fail(FVS_EINVAL, "fvs_mesh_cell_order: cell_offsets must increase (every instance has cells)");
      std::fill(inst.begin() + cell_offsets[k], inst.begin() + cell_offsets[k + 1], k);
    }
  }
  const std::vector<int> p = cell_order(m, reorder, inst);
  std::copy(p.begin(), p.end(), perm);
  return 0;
} catch (...) {
  return host_exception("fvs_mesh_cell_order");
}

int fvs_mesh_lsq_nnz(const fvs_mesh *M) { return (M && !M->mp.lsq_ptr.empty()) ? M->mp.lsq_ptr.back() : 0; }

int fvs_mesh_lsq(const fvs_mesh *M, int *ptr, int *cell, double *w, double *coef) try {
  if (!M) return fail(FVS_EINVAL, "null mesh");
  const MeshPrep &m = M->mp;
  if (m.lsq_ptr.empty()) return fail(FVS_ESTATE, "no LSQ stencil in this configuration");
  if (ptr) std::copy(m.lsq_ptr.begin(), m.lsq_ptr.end(), ptr);
  if (cell) std::copy(m.lsq_cell.begin(), m.lsq_cell.end(), cell);
  if (w) std::copy(m.lsq_w.begin(), m.lsq_w.end(), w);
  if (coef) std::copy(m.lsq_coef.begin(), m.lsq_coef.end(), coef);
  return 0;
} catch (...) {
  return host_exception("fvs_mesh_lsq");
}

int fvs_set_state(fvs_ctx *C, const double *cvar) try {
  if (!C || !cvar) return fail(FVS_EINVAL, "null argument");
  std::vector<double> h;
  to_device_order(*C, cvar, h, 4);
  HIPC(hipMemcpyAsync(C->U0, h.data(), h.size() * 8, hipMemcpyHostToDevice, C->stream));
  refresh_prim(*C);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(C->stream));
  C->have_state = true;
  return 0;
} catch (...) {
  return host_exception("fvs_set_state");
}

int fvs_get_state(fvs_ctx *C, double *cvar) try {
  if (int rc = need_state(C)) return rc;
  if (!cvar) return fail(FVS_EINVAL, "null argument");
  std::vector<double> h(4 * size_t(C->N));
  HIPC(hipMemcpyAsync(h.data(), C->U0, h.size() * 8, hipMemcpyDeviceToHost, C->stream));
  HIPC(hipStreamSynchronize(C->stream));
  to_ref_order(*C, h, cvar, 4);
  return 0;
} catch (...) {
  return host_exception("fvs_get_state");
}

int fvs_initialize(fvs_ctx *C, int mode, double t0) try {
  if (!C) return fail(FVS_EINVAL, "null context");
  // initial primitive state at the cell centres, then pvar2cvar
  // (src/initialize.f90:36-54,79-87; src/data_solution.f90:92-106)
  const MeshPrep &m = C->mp;
  const double kPi = 3.14159265358979323846, g = C->opt.gamma;
  std::vector<double> cv(4 * size_t(C->N));
  for (int c = 0; c < C->N; ++c) {
    double pv[4];
    const double x = m.cx[c], y = m.cy[c];
    if (mode == 0) {
      pv[0] = 1.12 + 0.15 * std::sin((3.12 * kPi) * x + (2.92 * kPi) * y);
      pv[1] = 1.32 + 0.06 * std::sin((2.09 * kPi) * x + (3.12 * kPi) * y);
      pv[2] = 1.18 + 0.03 * std::sin((2.15 * kPi) * x + (3.32 * kPi) * y);
      pv[3] = 1.62 + 0.31 * std::sin((3.79 * kPi) * x + (2.98 * kPi) * y);
    } else if (C->opt.lvortex) {
      const Phys &P = C->P;
      const double Tinf = P.vinf[3] / P.vinf[0];
      const double dx = x - (P.vx0 + P.vinf[1] * t0), dy = y - (P.vy0 + P.vinf[2] * t0);
      const double r = std::sqrt(dx * dx + dy * dy), cc = P.vK / (2.0 * kPi);
      pv[1] = P.vinf[1] - cc * dy * std::exp(0.5 * (1.0 - r * r));
      pv[2] = P.vinf[2] + cc * dx * std::exp(0.5 * (1.0 - r * r));
      const double temp = Tinf - (cc * cc) * (g - 1.0) / (2.0 * g) * std::exp(1.0 - r * r);
      pv[0] = std::pow(temp, 1.0 / (g - 1.0));
      pv[3] = std::pow(pv[0], g);
    } else {
      for (int k = 0; k < 4; ++k) pv[k] = C->P.pinf[k];
    }
    double *u = &cv[4 * size_t(c)];
    u[0] = pv[0];
    u[1] = pv[0] * pv[1];
    u[2] = pv[0] * pv[2];
    u[3] = pv[3] / (g - 1.0) + 0.5 * pv[0] * (pv[1] * pv[1] + pv[2] * pv[2]);
  }
  return fvs_set_state(C, cv.data());
} catch (...) {
  return host_exception("fvs_initialize");
}

int fvs_primitives(fvs_ctx *C, double *pvar) try {
  if (int rc = need_state(C)) return rc;
  refresh_prim(*C);
  HIPC(hipGetLastError());
  if (pvar) {
    std::vector<double> h(4 * size_t(C->N));
    HIPC(hipMemcpyAsync(h.data(), C->W, h.size() * 8, hipMemcpyDeviceToHost, C->stream));
    HIPC(hipStreamSynchronize(C->stream));
    to_ref_order(*C, h, pvar, 4);
  }
  HIPC(hipStreamSynchronize(C->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_primitives");
}

int fvs_gradient(fvs_ctx *C, double *grad) try {
  if (int rc = need_state(C)) return rc;
  ensure_prim(*C);
  if (C->opt.recon != 1)
    if (int rc = launch_gradient(*C, false)) return rc;
  if (grad) {
    std::vector<double> h(8 * size_t(C->N));
    HIPC(hipMemcpyAsync(h.data(), C->grad, h.size() * 8, hipMemcpyDeviceToHost, C->stream));
    HIPC(hipStreamSynchronize(C->stream));
    // device record [x0..x3, y0..y3] -> Fortran grad(4,N,2)
    const int N = C->N;
    for (int i = 0; i < N; ++i)
      for (int v = 0; v < 4; ++v) {
        grad[size_t(C->perm[i]) * 4 + v] = h[8 * size_t(i) + v];
        grad[size_t(N) * 4 + size_t(C->perm[i]) * 4 + v] = h[8 * size_t(i) + 4 + v];
      }
  }
  HIPC(hipStreamSynchronize(C->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_gradient");
}

int fvs_gradient_1var(fvs_ctx *C, const double *fc, double *dfc) try {
  if (!C || !fc || !dfc) return fail(FVS_EINVAL, "null argument");
  // the scalar rides in slot 0 of a scratch primitive record; the gradient
  // kernels treat the four slots independently
  const int N = C->N;
  std::vector<double> h(4 * size_t(N), 0.0);
  for (int i = 0; i < N; ++i) h[4 * size_t(i)] = fc[C->perm[i]];
  HIPC(hipMemcpyAsync(C->rout, h.data(), h.size() * 8, hipMemcpyHostToDevice, C->stream));
  std::swap(C->W, C->rout);
  int rc = launch_gradient(*C, false);
  std::swap(C->W, C->rout);
  if (rc) return rc;
  std::vector<double> g(8 * size_t(N));
  HIPC(hipMemcpyAsync(g.data(), C->grad, g.size() * 8, hipMemcpyDeviceToHost, C->stream));
  HIPC(hipStreamSynchronize(C->stream));
  for (int i = 0; i < N; ++i) {
    dfc[C->perm[i]] = g[8 * size_t(i)];
    dfc[size_t(N) + C->perm[i]] = g[8 * size_t(i) + 4];
  }
  return 0;
} catch (...) {
  return host_exception("fvs_gradient_1var");
}

int fvs_limiter(fvs_ctx *C, double *phi) try {
  if (int rc = need_state(C)) return rc;
  ensure_prim(*C);
  if (C->opt.recon != 1 && C->opt.limiter > 0)
    if (int rc = launch_gradient(*C, true)) return rc;
  if (phi) {
    std::vector<double> h(C->N);
    HIPC(hipMemcpyAsync(h.data(), C->phi, h.size() * 8, hipMemcpyDeviceToHost, C->stream));
    HIPC(hipStreamSynchronize(C->stream));
    to_ref_order(*C, h, phi, 1);
  }
  HIPC(hipStreamSynchronize(C->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_limiter");
}

int fvs_compute_residual(fvs_ctx *C, double time, double *resid, double *ws) try {
  if (int rc = need_state(C)) return rc;
  refresh_prim(*C);
  if (int rc = launch_fluxes(*C, time)) return rc;
  Upd A{};
  A.mode = M_RESID;
  A.gamma = C->opt.gamma;
  if (int rc = launch_update(*C, A)) return rc;
  std::vector<double> h;
  if (resid) {
    h.resize(4 * size_t(C->N));
    HIPC(hipMemcpyAsync(h.data(), C->rout, h.size() * 8, hipMemcpyDeviceToHost, C->stream));
    HIPC(hipStreamSynchronize(C->stream));
    to_ref_order(*C, h, resid, 4);
  }
  if (ws) {
    h.resize(C->N);
    HIPC(hipMemcpyAsync(h.data(), C->wsout, h.size() * 8, hipMemcpyDeviceToHost, C->stream));
    HIPC(hipStreamSynchronize(C->stream));
    to_ref_order(*C, h, ws, 1);
  }
  HIPC(hipStreamSynchronize(C->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_compute_residual");
}

int fvs_local_dt(fvs_ctx *C, double *dt) try {
  if (int rc = need_state(C)) return rc;
  // ws_nrml of the current state, then dt = CFL*vol/(0.5*ws)
  std::vector<double> ws(C->N);
  if (int rc = fvs_compute_residual(C, 0.0, nullptr, ws.data())) return rc;
  if (dt)
    for (int c = 0; c < C->N; ++c) dt[c] = C->opt.cfl * C->mp.vol[c] / (0.5 * ws[c]);
  return 0;
} catch (...) {
  return host_exception("fvs_local_dt");
}

int fvs_roe_flux_batch(int n, double gamma, const double *pL, const double *pR, const double *nxy, double *flux,
                       double *ws) try {
  if (n <= 0) return 0;
  if (!pL || !pR || !nxy || !flux || !ws) return fail(FVS_EINVAL, "null argument");
  double *d = nullptr;
  const size_t bytes = size_t(n) * 8 * (4 + 4 + 2 + 4 + 1);
  HIPC(hipMalloc(&d, bytes));
  double *dL = d, *dR = dL + 4 * size_t(n), *dn = dR + 4 * size_t(n), *dF = dn + 2 * size_t(n), *dw = dF + 4 * size_t(n);
  int rc = 0;
  if (hipMemcpy(dL, pL, 32 * size_t(n), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dR, pR, 32 * size_t(n), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dn, nxy, 16 * size_t(n), hipMemcpyHostToDevice) != hipSuccess) {
    rc = fail(FVS_EHIP, "hipMemcpy failed");
  } else {
    hipLaunchKernelGGL(k_roe_batch, dim3(nblk(n)), dim3(kBlock), 0, 0, n, gamma, dL, dR, dn, dF, dw);
    if (hipGetLastError() != hipSuccess || hipMemcpy(flux, dF, 32 * size_t(n), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ws, dw, 8 * size_t(n), hipMemcpyDeviceToHost) != hipSuccess)
      rc = fail(FVS_EHIP, "Roe batch kernel failed");
  }
  (void)hipFree(d);
  return rc;
} catch (...) {
  return host_exception("fvs_roe_flux_batch");
}

}  // extern "C"

namespace {

// The stage sequence of one time step (src/runge_kutta.f90:120-418 and the
// SSP-RK3 of this library).  Norm outputs go to dn / dinst at the step index
// held in the device counter, which the step advances; so the same launch
// sequence can be captured once in a HIP graph and replayed.
// SSP-RK3 stages on conservative states (no primitive-state traffic): the
// fused path with the staged-halo LSQ gradient, an unlimited Green-Gauss
// gradient (kernel_path 3 keeps primitive states) or a first-order
// reconstruction
bool cons_stages(const fvs_ctx &C) {
  // Green-Gauss gradients: unlimited, or limited with the bounds from the
  // staged halo (where the stage states are converted once, as staged)
  const bool gg = (C.opt.grad_method == 1 || C.opt.grad_method == 2) && C.opt.kernel_path != 3 &&
                  (C.opt.limiter == 0 || (C.halo && C.use_halo));
  return C.fused && !C.stage && (C.opt.rk_scheme == FVS_RK_SSP3 || C.opt.rk_scheme == FVS_RK_SSP42_SO) &&
         (C.opt.recon == 1 || (C.opt.grad_method == 3 && C.halo && C.use_halo) || gg);
}

// Update constants of stage k of a step starting at told, and the stage's
// time ts (runge_kutta_init, src/runge_kutta.f90:36-65; SSP-RK3 stages at
// t, t + dt, t + dt/2)
Upd stage_upd(const fvs_options &o, int k, double told, double &ts) {
  const double dt = o.dt;
  double ak = 1, bk = 2, ck = 2, dk = 1;  // runge_kutta_init, src/runge_kutta.f90:36-51
  if (o.rk_order == 1) ak = 3.60897, bk = 2.04, ck = 0.34206, dk = 0.00897;
  if (o.rk_order == 2) ak = 0.11, bk = 3.92, ck = 1.86, dk = 0.11;
  if (o.rk_order == 3) ak = 0.65, bk = 2.7, ck = 2.0, dk = 0.65;
  const double rk_coef[4] = {ak, bk, ck, dk};
  const double h_rk[4] = {dt / 2.0, dt / 2.0, dt, dt / 6.0};
  const double dts[4] = {0.0, dt / 3.0, dt * 2.0 / 3.0, dt};  // src/runge_kutta.f90:56-65
  const double ssp3_a[3] = {0.0, 0.75, 1.0 / 3.0}, ssp3_b[3] = {1.0, 0.25, 2.0 / 3.0};
  const int nst = (o.rk_scheme == FVS_RK_SSP3) ? 3 : 4;
  Upd A{};
  A.steady = o.steady;
  A.first = (k == 0);
  A.last = (k == nst - 1);
  A.cfl = o.cfl;
  A.gamma = o.gamma;
  ts = told;
  if (o.rk_scheme == FVS_RK_SSP3) {
    A.mode = M_SSP3;
    A.a = ssp3_a[k];
    A.b = ssp3_b[k];
    A.h = o.steady ? 1.0 : dt;
    ts = (k == 0) ? told : (k == 1 ? told + dt : told + 0.5 * dt);
  } else if (o.rk_scheme == FVS_RK_SSP42_SO) {
    // the SSP-RK(4,2) stages of the reference in two-register form, run by
    // the SSP-RK3 update: u_k = a u_0 + b (u_{k-1} + h R)
    const double a42[4] = {0.0, 0.0, 0.0, 0.25}, b42[4] = {1.0, 1.0, 1.0, 0.75};
    A.mode = M_SSP3;
    A.a = a42[k];
    A.b = b42[k];
    A.h = o.steady ? 1.0 / 3.0 : dt / 3.0;
    ts = told + dts[k];
  } else if (o.rk_scheme == FVS_RK_SSP42) {
    A.mode = M_SSP42;
    A.coef = 1.0;
    A.h = o.steady ? (k == 3 ? 1.0 / 4.0 : 1.0 / 3.0) : (k == 3 ? dt / 4.0 : dt / 3.0);
    ts = told + dts[k];
  } else {
    A.mode = M_CLASSIC;
    A.coef = rk_coef[k];
    A.use_acc = (k == 3);
    A.h = o.steady ? (k == 2 ? 1.0 : (k == 3 ? 1.0 / 6.0 : 1.0 / 2.0)) : h_rk[k];
    const double t12 = told + 0.5 * dt, tnew = told + dt;
    const double tst[4] = {told, t12, t12, tnew};
    ts = tst[k];
  }
  return A;
}

int launch_step(fvs_ctx &C, double told, double *dn, double *dinst) {
  const fvs_options &o = C.opt;
  const int nst = (o.rk_scheme == FVS_RK_SSP3) ? 3 : 4;
  for (int k = 0; k < nst; ++k) {
    double ts = told;
    const Upd A = stage_upd(o, k, told, ts);
    int rc;
    if (C.stage) {
      rc = launch_stage(C, A, ts);
    } else if (cons_stages(C)) {
      // the stage input and output states (a stage's output never
      // overwrites a state its neighbouring blocks still read)
      double *const in3[3] = {C.U0, C.Ub, C.Uc}, *const out3[3] = {C.Ub, C.Uc, C.U0};
      double *const in4[4] = {C.U0, C.Ub, C.Uc, C.Ub}, *const out4[4] = {C.Ub, C.Uc, C.Ub, C.U0};
      rc = nst == 3 ? launch_flux_update(C, A, ts, in3[k], out3[k]) : launch_flux_update(C, A, ts, in4[k], out4[k]);
    } else if (C.fused) {
      rc = launch_flux_update(C, A, ts);
    } else {
      rc = launch_fluxes(C, ts);
      if (!rc) rc = launch_update(C, A);
    }
    if (rc) return rc;
  }
  // the step counter (the norms' output row) advances after its last reader
  const bool inst = dinst && C.ninst > 0;
  if (!C.nact) {
    ProfScope ps(C, KC_NORM);
    // one partial sum per launch block of the last stage
    const int nparts = (C.fused && !C.stage && C.small_blocks) ? nblk(C.N, kSmallBlock) : nblk(C.N);
    hipLaunchKernelGGL(k_norm_final, dim3(1), dim3(kNormThreads), 0, C.stream, nparts, C.N, C.part, dn, C.dctr,
                       inst ? 0 : 1);
  }
  if (inst) {
    ProfScope ps(C, KC_NORM);
    hipLaunchKernelGGL(k_inst_norms, dim3(C.ninst), dim3(kBlock), 0, C.stream, C.ninst, C.dinst_off, C.d2buf, C.dctr,
                       dinst);
  }
  if (C.nact || inst) hipLaunchKernelGGL(k_counter, dim3(1), dim3(1), 0, C.stream, C.dctr, 1, 1);
  HIPC(hipGetLastError());
  return 0;
}

bool time_dependent(const fvs_ctx &C) {
  if (!C.opt.lvortex) return false;
  for (int b = 0; b < C.mp.nb; ++b)
    if (C.mp.bnd_type[b] == FVS_BC_DIRICHLET) return true;
  return false;
}

// capture gsteps steps into the context's graph, unless the cached one was
// captured with the same launches, pointers and grid sizes
int ensure_graph(fvs_ctx &C, int gsteps, double t1, double *dn, double *dinst) {
  fvs_ctx::GraphKey key;
  key.gen = C.launch_gen;
  key.gsteps = gsteps;
  key.dn = dn;
  key.di = dinst;
  key.w = C.W;
  key.walt = C.Walt;
  if (C.gexec && key == C.gkey) return 0;
  C.drop_graph();
  hipGraph_t graph = nullptr;
  HIPC(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  // t1 is baked into the captured launches; only time-dependent boundary
  // data read it, and those never run from a graph
  for (int g = 0; g < gsteps && !rc; ++g) rc = launch_step(C, t1, dn, dinst);
  hipError_t e = hipStreamEndCapture(C.stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(FVS_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(&C.gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    C.gexec = nullptr;
    return fail(FVS_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  }
  C.gkey = key;
  return 0;
}

// nsteps whole steps in one cooperative launch of k_steps_persistent (the
// whole-stage block body over the launch-block list, a grid barrier after
// every stage, the step norms in the same launch); swaps the primitive
// buffers like the per-stage launches would
int launch_persistent(fvs_ctx &C, double t1, int nsteps, double *dn, double *dinst) {
  const fvs_options &o = C.opt;
  PersistCtl c{};
  c.nst = (o.rk_scheme == FVS_RK_SSP3) ? 3 : 4;
  for (int k = 0; k < c.nst; ++k) {
    double ts = t1;
    c.A[k] = stage_upd(o, k, t1, ts);
  }
  c.nsteps = nsteps;
  c.nlaunch = C.nact ? C.nact : nblk(C.N);
  c.W0 = C.W;
  c.W1 = C.Walt;
  c.d2buf = C.d2buf;
  c.dn = C.nact ? nullptr : dn;
  c.nparts = nblk(C.N);
  const bool inst = dinst && C.ninst > 0;
  c.dinst = inst ? dinst : nullptr;
  c.ninst = C.ninst;
  c.inst_off = C.dinst_off;
  c.ctr = C.dctr;
  c.bar = C.pbar;
  c.err = C.perr;
  c.spin = 1u << 22;
  StageArgs a = stage_args(C, t1);
  a.nb = nblk(C.N);  // plain numbering: the persistent loop walks the blocks itself
  const size_t lds = std::max(stage_lds(C), size_t(4) * kNormThreads * sizeof(double));
  const bool lim = o.limiter > 0 && !C.limiter_frozen;
  const void *fn = nullptr;
#define FVS_PK(L, PM) fn = reinterpret_cast<const void *>(&k_steps_persistent<L, PM>)
  if (lim) {
    if (C.pmode == 0) FVS_PK(1, 0);
    else if (C.pmode == 1) FVS_PK(1, 1);
    else if (C.pmode == 2) FVS_PK(1, 2);
    else FVS_PK(1, 3);
  } else {
    if (C.pmode == 0) FVS_PK(0, 0);
    else if (C.pmode == 1) FVS_PK(0, 1);
    else if (C.pmode == 2) FVS_PK(0, 2);
    else FVS_PK(0, 3);
  }
#undef FVS_PK
  int per_cu = 0;
  HIPC(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kStageThreads, lds));
  if (per_cu < 1) return fail(FVS_ESTATE, "persistent launch: the whole-stage kernel does not fit on a compute unit");
  const int grid = std::min(c.nlaunch, per_cu * C.ncu);
  hipLaunchKernelGGL(k_barrier_reset, dim3(1), dim3(64), 0, C.stream, C.pbar, C.perr);
  void *args[] = {&a, &c};
  HIPC(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kStageThreads), args, unsigned(lds), C.stream));
  C.persist_ran = true;
  if ((nsteps * c.nst) % 2) std::swap(C.W, C.Walt);
  return 0;
}

int run_steps(fvs_ctx &C, double t1, int nsteps, double *dn, double *dv, double *dinst, bool graph_ok = true) {
  const double dt = C.opt.dt;
  const int nst = (C.opt.rk_scheme == FVS_RK_SSP3) ? 3 : 4;
  const bool cons = cons_stages(C);
  if (!cons) ensure_prim(C);  // the stages read W
  hipLaunchKernelGGL(k_counter, dim3(1), dim3(1), 0, C.stream, C.dctr, 0, 0);
  HIPC(hipGetLastError());
  if (graph_ok && C.use_persistent && C.stage && !C.profiling && !dv && !time_dependent(C))
    return launch_persistent(C, t1, nsteps, dn, dinst);
  int s = 0;
  // graph replay: one captured graph covers an even number of stages so the
  // primitive-state buffers are back in place after each replay
  if (graph_ok && C.use_graph && !C.profiling && !dv && !time_dependent(C)) {
    const int gsteps = (nst % 2 == 0) ? 1 : 2;
    if (nsteps >= 2 * gsteps) {
      if (int rc = ensure_graph(C, gsteps, t1, dn, dinst)) return rc;
      hipError_t e = hipSuccess;
      for (; s + gsteps <= nsteps && e == hipSuccess; s += gsteps) e = hipGraphLaunch(C.gexec, C.stream);
      if (e != hipSuccess) return fail(FVS_EHIP, std::string("graph launch: ") + hipGetErrorString(e));
    }
  }
  for (; s < nsteps; ++s) {
    const double told = t1 + double(s) * dt;
    if (int rc = launch_step(C, told, dn, dinst)) return rc;
    if (dv)
      if (int rc = launch_vortex_err(C, told + dt, dv + kVerr * size_t(s))) return rc;
  }
  // the conservative stages wrote no primitive state: W is refreshed by the
  // next caller that reads it, not inside the timed steps
  if (cons) C.w_stale = true;
  return 0;
}

int integrate(fvs_ctx *C, double t1, int nsteps, double *norms, double *verr, double *inst_norms) {
  if (int rc = need_state(C)) return rc;
  if (nsteps < 0) return fail(FVS_EINVAL, "nsteps < 0");
  if (nsteps == 0) return 0;
  // device outputs: the context's buffers (allocated only when a call needs
  // more than any earlier one)
  if (int rc = C->grow(C->dn_buf, C->dn_cap, size_t(nsteps) * 4)) return rc;
  if (verr)
    if (int rc = C->grow(C->dv_buf, C->dv_cap, size_t(nsteps) * kVerr)) return rc;
  const bool want_inst = inst_norms && C->ninst > 0;
  if (want_inst)
    if (int rc = C->grow(C->di_buf, C->di_cap, size_t(nsteps) * C->ninst * 4)) return rc;
  double *dn = C->dn_buf, *dv = verr ? C->dv_buf : nullptr, *di = want_inst ? C->di_buf : nullptr;
  const bool timers = C->phase_timing && !C->profiling;
  if (timers) {
    C->profiling = true;
    C->evused = 0;
    C->evclass.clear();
  }
  HIPC(hipEventRecord(C->ev0, C->stream));
  const int rc_run = run_steps(*C, t1, nsteps, dn, dv, di);
  if (timers) C->profiling = false;
  if (rc_run) return rc_run;
  HIPC(hipEventRecord(C->ev1, C->stream));
  // the small outputs queue behind the steps; one wait for all of them
  int perr = 0;
  if (C->persist_ran) HIPC(hipMemcpyAsync(&perr, C->perr, sizeof(int), hipMemcpyDeviceToHost, C->stream));
  if (norms) HIPC(hipMemcpyAsync(norms, dn, size_t(nsteps) * 4 * 8, hipMemcpyDeviceToHost, C->stream));
  if (dv) HIPC(hipMemcpyAsync(verr, dv, size_t(nsteps) * kVerr * 8, hipMemcpyDeviceToHost, C->stream));
  if (di) HIPC(hipMemcpyAsync(inst_norms, di, size_t(nsteps) * C->ninst * 4 * 8, hipMemcpyDeviceToHost, C->stream));
  HIPC(hipStreamSynchronize(C->stream));
  if (C->persist_ran) {
    C->persist_ran = false;
    if (perr) return fail(FVS_EHIP, "persistent launch: a workgroup gave up waiting at a grid barrier");
  }
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, C->ev0, C->ev1));
  C->last_ms = ms;
  if (timers)
    for (size_t k = 0; k < C->evclass.size(); ++k) {
      float t = 0.f;
      HIPC(hipEventElapsedTime(&t, C->evpool[2 * k], C->evpool[2 * k + 1]));
      C->phase_ms[C->evclass[k]] += t;
    }
  return 0;
}

}  // namespace

extern "C" {

int fvs_time_integration(fvs_ctx *C, double t1, int nsteps, double *norms, double *verr) try {
  if (C && C->nact && norms) {
    // with inactive instances the whole-mesh norm would mix stale partials
    std::fill(norms, norms + 4 * size_t(std::max(nsteps, 0)), std::nan(""));
    return integrate(C, t1, nsteps, nullptr, verr, nullptr);
  }
  return integrate(C, t1, nsteps, norms, verr, nullptr);
} catch (...) {
  return host_exception("fvs_time_integration");
}

int fvs_time_integration_noalloc(fvs_ctx *C, double t1, int nsteps, double *norms, double *verr) try {
  if (int rc = need_state(C)) return rc;
  if (nsteps < 0) return fail(FVS_EINVAL, "nsteps < 0");
  if (nsteps > kNoAllocMaxSteps)
    return fail(FVS_EINVAL, "fvs_time_integration_noalloc: nsteps exceeds the " + std::to_string(kNoAllocMaxSteps) +
                                " steps its device staging holds; call it in chunks");
  if (nsteps == 0) return 0;
  if (!C->scratch_dev) return fail(FVS_ESTATE, "no device staging address for this context");
  if (C->nact && norms) {
    std::fill(norms, norms + 4 * size_t(nsteps), std::nan(""));
    norms = nullptr;
  }
  // staging in the device's `scratch`: norms of step s at [4 s, 4 s + 4),
  // vortex errors after all the norms; every value is written by the steps
  // before it is copied out, so nothing needs clearing between calls
  double *dn = C->scratch_dev, *dv = verr ? C->scratch_dev + 4 * size_t(nsteps) : nullptr;
  HIPC(hipEventRecord(C->ev0, C->stream));
  if (int rc = run_steps(*C, t1, nsteps, dn, dv, nullptr, /*graph_ok=*/false)) return rc;
  HIPC(hipEventRecord(C->ev1, C->stream));
  if (norms) HIPC(hipMemcpyAsync(norms, dn, size_t(nsteps) * 4 * 8, hipMemcpyDeviceToHost, C->stream));
  if (dv) HIPC(hipMemcpyAsync(verr, dv, size_t(nsteps) * kVerr * 8, hipMemcpyDeviceToHost, C->stream));
  HIPC(hipStreamSynchronize(C->stream));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, C->ev0, C->ev1));
  C->last_ms = ms;
  return 0;
} catch (...) {
  return host_exception("fvs_time_integration_noalloc");
}

int fvs_set_phase_timers(fvs_ctx *C, int on) try {
  if (!C) return fail(FVS_EINVAL, "null context");
  C->phase_timing = on != 0;
  return 0;
} catch (...) {
  return host_exception("fvs_set_phase_timers");
}

int fvs_phase_times(const fvs_ctx *C, double *sec4) try {
  if (!C || !sec4) return fail(FVS_EINVAL, "null argument");
  const double *t = C->phase_ms;
  sec4[0] = t[KC_GRAD] * 1e-3;                                    // gradient, limiter fused in
  sec4[1] = 0.0;                                                  // limiter: inside the gradient launches
  sec4[2] = (t[KC_FLUX_INT] + t[KC_FLUX_BND] + t[KC_FUSED]) * 1e-3;  // face fluxes (+ fused stage update)
  sec4[3] = (t[KC_UPDATE] + t[KC_NORM]) * 1e-3;                   // stage updates, residual norms
  return 0;
} catch (...) {
  return host_exception("fvs_phase_times");
}

int fvs_set_persistent(fvs_ctx *C, int on) try {
  if (!C) return fail(FVS_EINVAL, "null context");
  if (on && !C->stage)
    return fail(FVS_ESTATE, "persistent launches run the whole-stage kernel: create the context with kernel_path 2 "
                            "(LSQ gradient, triangle-sized blocks)");
  C->use_persistent = on != 0;
  return 0;
} catch (...) {
  return host_exception("fvs_set_persistent");
}

int fvs_set_graph(fvs_ctx *C, int on) try {
  if (!C) return fail(FVS_EINVAL, "null context");
  C->use_graph = on != 0;
  if (!C->use_graph) C->drop_graph();
  return 0;
} catch (...) {
  return host_exception("fvs_set_graph");
}

int fvs_batch_time_integration(fvs_ctx *C, double t1, int nsteps, double *inst_norms) try {
  if (!C) return fail(FVS_EINVAL, "null context");
  if (C->ninst <= 0) return fail(FVS_ESTATE, "not a batch context (use fvs_create_batch)");
  return integrate(C, t1, nsteps, nullptr, nullptr, inst_norms);
} catch (...) {
  return host_exception("fvs_batch_time_integration");
}

int fvs_batch_set_active(fvs_ctx *C, const int *active) try {
  if (!C || !active) return fail(FVS_EINVAL, "null argument");
  if (C->ninst <= 0) return fail(FVS_ESTATE, "not a batch context (use fvs_create_batch)");
  if (!C->fused || !(C->lsqc || C->opt.recon == 1))
    return fail(FVS_ESTATE, "instance masking needs the fused kernels and the centroid-form LSQ gradient");
  const int nb = nblk(C->N);
  std::vector<int> list;
  bool all = true;
  for (int k = 0; k < C->ninst; ++k) all = all && active[k];
  if (!all) {
    // launch the blocks that hold a cell of an active instance; a block that
    // also holds masked cells leaves them alone (its mask bits are clear)
    std::vector<unsigned> mask(8 * size_t(nb), 0u);
    const std::vector<int> &off = C->inst_off_dev;
    for (int b = 0; b < nb; ++b) {
      const int i0 = b * kBlock, i1 = std::min(C->N, i0 + kBlock);
      const int k0 = int(std::upper_bound(off.begin(), off.end(), i0) - off.begin()) - 1;
      const int k1 = int(std::upper_bound(off.begin(), off.end(), i1 - 1) - off.begin()) - 1;
      bool on = false;
      for (int k = k0; k <= k1; ++k) {
        if (!active[k]) continue;
        on = true;
        for (int i = std::max(off[k], i0); i < std::min(off[k + 1], i1); ++i)
          mask[8 * size_t(b) + (i - i0) / 32] |= 1u << ((i - i0) % 32);
      }
      if (on) list.push_back(b);
    }
    if (list.empty()) list.push_back(0);
    HIPC(hipMemcpyAsync(C->blist, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, C->stream));
    HIPC(hipMemcpyAsync(C->blist + nb, mask.data(), mask.size() * sizeof(unsigned), hipMemcpyHostToDevice, C->stream));
  }
  // masked cells are not updated, so on the primitive-state paths their W
  // lags the stage parity; a resumed instance starts from its own state
  refresh_prim(*C);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(C->stream));
  C->nact = all ? 0 : int(list.size());
  C->launch_gen++;
  return 0;
} catch (...) {
  return host_exception("fvs_batch_set_active");
}

int fvs_batch_wall_forces(fvs_ctx *C, double *out) try {
  if (int rc = need_state(C)) return rc;
  if (C->ninst <= 0) return fail(FVS_ESTATE, "not a batch context (use fvs_create_batch)");
  if (!out) return fail(FVS_EINVAL, "null argument");
  refresh_prim(*C);
  const int save = C->nact;
  C->nact = 0;
  int rc = launch_gradient(*C, false);
  C->nact = save;
  if (rc) return rc;
  hipLaunchKernelGGL(k_wall_forces_inst, dim3(nblk(C->ninst)), dim3(kBlock), 0, C->stream, C->ninst, C->dwoff,
                     C->wall, C->fb_cell, C->fb_geo, C->W, C->grad, C->dmach_aoa, C->dpinf_inst, C->dforces);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(out, C->dforces, 4 * size_t(C->ninst) * 8, hipMemcpyDeviceToHost, C->stream));
  HIPC(hipStreamSynchronize(C->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_batch_wall_forces");
}

int fvs_vortex_error(fvs_ctx *C, double t, double *out15) try {
  if (int rc = need_state(C)) return rc;
  if (!out15) return fail(FVS_EINVAL, "null argument");
  if (int rc = launch_vortex_err(*C, t, C->small)) return rc;
  HIPC(hipMemcpyAsync(out15, C->small, kVerr * 8, hipMemcpyDeviceToHost, C->stream));
  HIPC(hipStreamSynchronize(C->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_vortex_error");
}

int fvs_wall_forces(fvs_ctx *C, double *out4) try {
  if (int rc = need_state(C)) return rc;
  if (!out4) return fail(FVS_EINVAL, "null argument");
  const double kPi = 3.14159265358979323846;
  refresh_prim(*C);
  if (int rc = launch_gradient(*C, false)) return rc;
  const double al = C->opt.aoa_deg * (kPi / 180.0);
  hipLaunchKernelGGL(k_wall_forces, dim3(1), dim3(64), 0, C->stream, C->nwall, C->wall, C->fb_cell, C->fb_geo, C->W,
                     C->grad, C->opt.mach, C->P.pinf[3], std::cos(al), std::sin(al), C->small + 16);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(out4, C->small + 16, 4 * 8, hipMemcpyDeviceToHost, C->stream));
  HIPC(hipStreamSynchronize(C->stream));
  return 0;
} catch (...) {
  return host_exception("fvs_wall_forces");
}

int fvs_wall_cp(fvs_ctx *C, int *nwall, double *out4, double *un3) try {
  if (int rc = need_state(C)) return rc;
  if (!nwall) return fail(FVS_EINVAL, "null argument");
  *nwall = C->nwall;
  if (un3) un3[0] = un3[1] = un3[2] = 0.0;
  if (C->nwall == 0 || (!out4 && !un3)) return 0;
  refresh_prim(*C);
  const int save = C->nact;
  C->nact = 0;
  int rc = launch_gradient(*C, false);
  C->nact = save;
  if (rc) return rc;
  const size_t n4 = 4 * size_t(C->nwall);
  double *d = nullptr;
  HIPC(hipMalloc(&d, n4 * sizeof(double)));
  const bool batch = C->ninst > 0;
  hipLaunchKernelGGL(k_wall_cp, dim3(nblk(C->nwall)), dim3(kBlock), 0, C->stream, C->nwall, C->wall, C->fb_cell,
                     C->fb_geo, C->W, C->grad, C->opt.mach, C->P.pinf[3], batch ? C->dbinst : nullptr,
                     batch ? C->dmach_aoa : nullptr, batch ? C->dpinf_inst : nullptr, d);
  std::vector<double> h(n4);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, n4 * sizeof(double), hipMemcpyDeviceToHost, C->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(C->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(FVS_EHIP, std::string("wall cp: ") + hipGetErrorString(e));
  if (out4) std::copy(h.begin(), h.end(), out4);
  if (un3) {
    // the reference's loop order and arithmetic (src/io.f90:413-422)
    double mx = 0.0, l2 = 0.0, l1 = 0.0;
    for (int k = 0; k < C->nwall; ++k) {
      const double un = h[4 * size_t(k) + 3];
      l2 = l2 + un * un;
      l1 = l1 + std::fabs(un);
      mx = std::max(mx, std::fabs(un));
    }
    un3[0] = mx;
    un3[1] = std::sqrt(l2 / double(C->nwall));
    un3[2] = l1 / double(C->nwall);
  }
  return 0;
} catch (...) {
  return host_exception("fvs_wall_cp");
}

int fvs_set_timestep(fvs_ctx *C, double dt, double cfl) try {
  if (!C) return fail(FVS_EINVAL, "null context");
  if (!(dt > 0.0) || !(cfl > 0.0)) return fail(FVS_EINVAL, "dt and cfl must be positive");
  C->opt.dt = dt;
  C->opt.cfl = cfl;
  C->launch_gen++;
  return 0;
} catch (...) {
  return host_exception("fvs_set_timestep");
}

int fvs_kernel_times(fvs_ctx *C, double t1, int nsteps, double *ms, double *bytes, int *launches) try {
  if (int rc = need_state(C)) return rc;
  if (nsteps <= 0 || !ms || !bytes || !launches) return fail(FVS_EINVAL, "bad argument");
  C->profiling = true;
  C->evused = 0;
  C->evclass.clear();
  int rc = fvs_time_integration(C, t1, nsteps, nullptr, nullptr);
  C->profiling = false;
  if (rc) return rc;
  double tot[KC_N] = {0, 0, 0, 0, 0, 0};
  int cnt[KC_N] = {0, 0, 0, 0, 0, 0};
  for (size_t k = 0; k < C->evclass.size(); ++k) {
    float t = 0.f;
    HIPC(hipEventElapsedTime(&t, C->evpool[2 * k], C->evpool[2 * k + 1]));
    tot[C->evclass[k]] += t;
    cnt[C->evclass[k]]++;
  }
  // algorithmic bytes per launch: every array the kernel must stream, each
  // touched once; neighbour re-reads of W / grad / phi / FA are cache hits
  const MeshPrep &m = C->mp;
  const double N = C->N, Fi = C->Fi, Fb = C->Fb;
  const bool lim = C->opt.limiter > 0 && C->opt.recon != 1;
  double bg = 0.0;
  if (C->opt.grad_method == 3 && C->halo && C->use_halo)
    bg = N * (32 + 16 + 24 + 64) + 4.0 * double(C->nghl) + 4.0 * double(C->nlcode);
  else if (C->lsqc) bg = N * (32 + 16 + 24 + 64) + 4.0 * m.lsq_ptr.back();
  else if (C->opt.grad_method == 3) bg = N * (32 + 64) + 20.0 * m.lsq_ptr.back();
  else if (C->opt.grad_method == 1) bg = N * (32 + 16 + 8 + 64) + N * C->NFW * 20.0;
  // GGNB: pass 1 reads the node incidences and the states, writes F; pass 2
  // reads its node slots and coefedg, F, vol, writes grad
  else bg = 12.0 * double(C->nnz_node) + N * 32 + 2.0 * 32 * C->NV + N * (20.0 * C->gwn + 8 + 64);
  if (lim) {
    bg += N * (8 + 4) + 16.0 * (3 * m.ntri + 4 * m.nquad);
    if (C->opt.grad_method == 3) bg += 0.0;
    else if (C->halo && C->use_halo) bg += 4.0 * double(C->nghl) + 4.0 * double(C->nlcode);  // staged halo
    else bg += 4.0 * m.lsq_ptr.back() + 8.0 * (C->opt.grad_method == 1 ? N : 0.0);
  }
  const double bfi = Fi * (8 + 56 + 40) + N * (32 + 64 + 8);
  const double bfb = Fb * (8 + 56 + 40) + Fb * (32 + 64 + 8);
  // update: gather ids, face fluxes once, vol, W write, state traffic per stage
  double state = 0.0;
  const int nst = (C->opt.rk_scheme == FVS_RK_SSP3) ? 3 : 4;
  if (C->opt.rk_scheme == FVS_RK_SSP3) state = (64 + 96 + 96) / 3.0;
  else if (C->opt.rk_scheme == FVS_RK_SSP42) state = (32 + 32 + 64 + 64 + 64 + 64 + 32 + 32 + 32) / 4.0;
  else state = (32 + 32 + 64 + 64 + 64 + 64 + 32 + 32 + 32) / 4.0;
  if (C->opt.steady) state += (8.0 + 8.0 * (nst - 1)) / nst;
  const double bup = N * (C->GW * 4.0 + 8 + 32 + state) + (Fi + Fb) * 40.0;
  const double bno = 4.0 * 8 * nblk(C->N) + 32;
  // fused: the listed face records (cells + 7 geometry values; a face on
  // a block border is listed, and stored, once per block), the stage's W and
  // grad once (phi unless there is no limiter), the cell-side traffic of the
  // update kernel without the face-flux buffer
  const bool phi1 = C->opt.limiter == 0 && C->opt.recon != 1;
  const double nlisted = double(C->small_blocks ? C->nlisted_s : C->nlisted);
  double bfu = nlisted * (8 + 56) + N * (32 + 64 + (phi1 ? 0 : 8)) + N * (C->GW * 2.0 + 8 + 32 + state);
  if (cons_stages(*C)) {
    // conservative stages: the stage input (read once, also the update's
    // stage state; on the first stage also its step-start state), the
    // step-start state on the other stages, the new state; no W write
    double st3 = (32 + 64 + 64) / 3.0;
    // two-register SSP-RK(4,2): the step-start state is read on the last
    // stage only
    if (C->opt.rk_scheme == FVS_RK_SSP42_SO) st3 = (32 + 32 + 32 + 64) / 4.0;
    if (C->opt.steady) st3 += (8.0 + 8.0 * (nst - 1)) / nst;
    bfu = nlisted * (8 + 56) + N * (32 + 64 + (phi1 ? 0 : 8)) + N * (C->GW * 2.0 + 8 + st3);
  }
  // whole stage: own states, centres and G^-1 once, the per-block cell lists
  // and local stencils, face records and block lists once, limiter inputs,
  // the update's cell-side traffic
  const double bst = N * (32 + 16 + 24) + 2.0 * double(C->nslid) + 4.0 * double(C->nlist2) + (Fi + Fb) * 56 +
                     8.0 * double(C->nlisted) + (lim ? N * (8 + 4) + 16.0 * (3 * m.ntri + 4 * m.nquad) : 0.0) +
                     N * (C->GW * 2.0 + 8 + 32 + state);
  const double b5[KC_N] = {bg, bfi, bfb, bup, bno, C->stage ? bst : bfu};
  for (int k = 0; k < KC_N; ++k) {
    ms[k] = cnt[k] ? tot[k] / cnt[k] : 0.0;
    bytes[k] = b5[k];
    launches[k] = cnt[k];
  }
  return 0;
} catch (...) {
  return host_exception("fvs_kernel_times");
}

int fvs_freeze_limiter(fvs_ctx *C, int frozen) try {
  if (!C) return fail(FVS_EINVAL, "null context");
  C->limiter_frozen = frozen != 0;
  C->frz_some = false;
  C->launch_gen++;
  return 0;
} catch (...) {
  return host_exception("fvs_freeze_limiter");
}

int fvs_batch_freeze_limiter(fvs_ctx *C, const int *frozen) try {
  if (!C || !frozen) return fail(FVS_EINVAL, "null argument");
  if (C->ninst <= 0) return fail(FVS_ESTATE, "not a batch context (use fvs_create_batch)");
  if (!C->lsqc) return fail(FVS_ESTATE, "per-instance freezing needs the centroid-form LSQ gradient");
  int nf = 0;
  for (int k = 0; k < C->ninst; ++k) nf += frozen[k] != 0;
  C->limiter_frozen = nf == C->ninst;
  C->frz_some = nf > 0 && nf < C->ninst;
  C->launch_gen++;
  if (C->frz_some) {
    if (!C->dfrz)
      if (int rc = C->alloc(C->dfrz, size_t(C->ninst))) return rc;
    std::vector<int> f(frozen, frozen + C->ninst);
    for (int &v : f) v = v != 0;
    HIPC(hipMemcpyAsync(C->dfrz, f.data(), f.size() * sizeof(int), hipMemcpyHostToDevice, C->stream));
    HIPC(hipStreamSynchronize(C->stream));
  }
  return 0;
} catch (...) {
  return host_exception("fvs_batch_freeze_limiter");
}

int fvs_autotune(fvs_ctx *C, int nsteps, int *choice4, double *best_ms) try {
  if (int rc = need_state(C)) return rc;
  if (nsteps <= 0) return fail(FVS_EINVAL, "nsteps must be positive");
  // the trials advance the state: keep a copy and restore it afterwards
  const size_t n4 = 4 * size_t(C->N);
  double *save = nullptr;
  HIPC(hipMalloc(&save, n4 * sizeof(double)));
  int rc = 0;
  if (hipMemcpyAsync(save, C->U0, n4 * 8, hipMemcpyDeviceToDevice, C->stream) != hipSuccess)
    rc = fail(FVS_EHIP, "autotune: state copy failed");
  auto trial = [&](double &ms) -> int {
    if (hipMemcpyAsync(C->U0, save, n4 * 8, hipMemcpyDeviceToDevice, C->stream) != hipSuccess)
      return fail(FVS_EHIP, "autotune: state restore failed");
    refresh_prim(*C);
    C->launch_gen++;  // the launch choices may have changed
    if (int r = fvs_time_integration(C, 0.0, nsteps, nullptr, nullptr)) return r;
    ms = C->last_ms / nsteps;
    return 0;
  };
  bool *knobs[7] = {&C->xcd_grad, &C->xcd_flux,  &C->xcd_upd,   &C->xcd_fused,
                    &C->use_halo, &C->small_blocks, &C->small_grad};
  const bool lsq_halo = C->halo && !C->stage && C->opt.recon != 1 && C->opt.grad_method == 3;
  const bool used[7] = {C->opt.recon != 1, !C->fused, !C->fused, C->fused && !C->stage,
                        C->halo && !C->stage && C->opt.recon != 1, C->have_small && C->fused && !C->stage,
                        lsq_halo && C->have_small_grad};
  double best = 0.0;
  // warm-up: the first few dozen launches after the device was idle run
  // slower (measured up to 15% for the fused flux kernel)
  for (int w = 0; w < 3 && !rc; ++w) rc = trial(best);
  for (int k = 0; k < 7 && !rc; ++k) {
    if (!used[k]) continue;
    if (k == 6 && !C->use_halo) continue;  // 128-cell blocks exist for the staged-halo gradient only
    // the current and the flipped setting alternately, twice each, so that a
    // drift of the device's speed during the search does not decide
    double cur = 0.0, flip = 0.0;
    for (int r = 0; r < 2 && !rc; ++r) {
      double a = 0.0, b = 0.0;
      rc = trial(a);
      *knobs[k] = !*knobs[k];
      if (!rc) rc = trial(b);
      *knobs[k] = !*knobs[k];
      cur += a;
      flip += b;
    }
    if (rc) break;
    if (flip < cur) {
      *knobs[k] = !*knobs[k];
      best = 0.5 * flip;
    } else {
      best = 0.5 * cur;
    }
  }
  if (hipMemcpyAsync(C->U0, save, n4 * 8, hipMemcpyDeviceToDevice, C->stream) != hipSuccess && !rc)
    rc = fail(FVS_EHIP, "autotune: state restore failed");
  refresh_prim(*C);
  if (hipStreamSynchronize(C->stream) != hipSuccess && !rc) rc = fail(FVS_EHIP, "autotune: sync failed");
  (void)hipFree(save);
  C->launch_gen++;
  if (rc) return rc;
  if (choice4) {
    for (int k = 0; k < 4; ++k) choice4[k] = *knobs[k] ? 1 : 0;
    if (C->halo && C->use_halo) choice4[0] |= 2;
    if (C->halo && C->use_halo && C->have_small_grad && C->small_grad) choice4[0] |= 4;
    if (C->small_blocks) choice4[3] |= 4;
  }
  if (best_ms) *best_ms = best;
  return 0;
} catch (...) {
  return host_exception("fvs_autotune");
}

int fvs_set_launch_choices(fvs_ctx *C, const int *choice4) try {
  if (!C || !choice4) return fail(FVS_EINVAL, "null argument");
  const bool halo = (choice4[0] & 2) != 0, small = (choice4[3] & 4) != 0, small_grad = (choice4[0] & 4) != 0;
  if (halo && !C->halo) return fail(FVS_EINVAL, "launch choices: no staged-halo gradient data for this context");
  if (small_grad && !(halo && C->have_small_grad))
    return fail(FVS_EINVAL, "launch choices: no 128-cell staged-halo gradient lists for this context");
  if (small && !(C->have_small && C->fused && !C->stage))
    return fail(FVS_EINVAL, "launch choices: no 128-cell block lists for this context");
  C->xcd_grad = (choice4[0] & 1) != 0;
  C->xcd_flux = (choice4[1] & 1) != 0;
  C->xcd_upd = (choice4[2] & 1) != 0;
  C->xcd_fused = (choice4[3] & 1) != 0;
  C->use_halo = halo;
  C->small_grad = small_grad;
  C->small_blocks = small;
  C->launch_gen++;
  return 0;
} catch (...) {
  return host_exception("fvs_set_launch_choices");
}

double fvs_last_elapsed_ms(const fvs_ctx *C) { return C ? C->last_ms : 0.0; }

}  // extern "C"
