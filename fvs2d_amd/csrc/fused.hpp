// Bandwidth-lean kernels of the time loop (included by fvs_api.hip inside its
// anonymous namespace, after the generic kernels and helpers).
//
// * XCD-aware block numbering: MI355X hands consecutive workgroups to its 8
//   XCDs round robin; remapping makes each XCD walk one contiguous eighth of
//   the cells, so the neighbour gathers of a block hit the L2 of its own XCD.
// * k_grad_lsqc: the LSQ gradient as G^-1 * sum_j w_j^2 d_j (q_j - q_i), with
//   d_j from the cell centres and G^-1 stored per cell, so a stencil entry
//   costs 4 bytes (the neighbour id) instead of 20 (id + two coefficients).
//   Mathematically the reference's coef_j w_j (src/gradient_lsq.f90:397-398);
//   the rounding differs in the last bits.
// * k_flux_update: one workgroup owns a block of 256 consecutive cells.  It
//   first computes the Roe flux of every face touching the block into LDS
//   (faces shared with another block are computed by both), then each thread
//   gathers its cell's faces from LDS in the reference's summation order and
//   applies the stage update.  Face fluxes never travel through HBM.

struct ShiftTable {
  double sx[15], sy[15];
};

// One stencil term of the LSQ gradient, shared by every gradient kernel
// (k_grad_lsqc, k_grad_lsqh, k_stage): neighbour centre (xj, yj) and state
// q, cell centre ci and state w.  The multiply-adds are written as explicit
// FMAs, so the rounding is the same in every kernel whatever the compiler's
// contraction choices in each of them (the kernels stay bit-identical).
template <int LIM, int PMODE>
__device__ __forceinline__ void lsq_term(double xj, double yj, const double2 ci, const double q[4],
                                         const double w[4], double p, double sx[4], double sy[4], double mn[4],
                                         double mx[4]) {
  const double dx = xj - ci.x, dy = yj - ci.y;
  double w2 = 1.0;
  if (PMODE != 0) {
    const double r2 = fma(dx, dx, dy * dy);
    if (PMODE == 1) w2 = (r2 > 0.0) ? 1.0 / r2 : 0.0;
    else if (PMODE == 2) w2 = (r2 > 0.0) ? 1.0 / (r2 * r2) : 0.0;
    else w2 = (r2 > 0.0) ? pow(r2, -p) : 0.0;
  }
  for (int v = 0; v < 4; ++v) {
    const double t = w2 * (q[v] - w[v]);
    sx[v] = fma(t, dx, sx[v]);
    sy[v] = fma(t, dy, sy[v]);
    if (LIM) {
      mn[v] = (q[v] < mn[v]) ? q[v] : mn[v];
      mx[v] = (q[v] > mx[v]) ? q[v] : mx[v];
    }
  }
}

// gradient = G^-1 (sx, sy) with G^-1's entries 11, 12, 22 (explicit FMAs as above)
__device__ __forceinline__ void lsq_apply_ginv(double G11, double G12, double G22, const double sx[4],
                                               const double sy[4], double gx[4], double gy[4]) {
  for (int v = 0; v < 4; ++v) {
    gx[v] = fma(G11, sx[v], G12 * sy[v]);
    gy[v] = fma(G12, sx[v], G22 * sy[v]);
  }
}

// PMODE: 0 -> p = 0, 1 -> p = 1, 2 -> p = 2, 3 -> general p (pow)
template <int LIM, int PMODE>
__global__ __launch_bounds__(kBlock) void k_grad_lsqc(int N, int nb, const double *__restrict__ W,
                                                      const double2 *__restrict__ cxy,
                                                      const double *__restrict__ ginv, const int *__restrict__ soff,
                                                      const int *__restrict__ swid, const unsigned *__restrict__ sid,
                                                      ShiftTable st, double p, double *__restrict__ grad,
                                                      double *__restrict__ phi, int lt, const int *__restrict__ nvrt,
                                                      const double2 *__restrict__ loff,
                                                      const double *__restrict__ vol,
                                                      const int *__restrict__ blist,
                                                      const int *__restrict__ ioff,
                                                      const int *__restrict__ frz, int ninst) {
  const int lb = blist ? blist[blockIdx.x] : xcd_block(nb);
  if (lb >= (nb < 0 ? -nb : nb)) return;  // the whole workgroup leaves before the barrier
  // the block's own states and centres in LDS: neighbours inside the block
  // (most of them, in Morton order) are read from there
  __shared__ double sW[4 * kBlock];
  __shared__ double2 sX[kBlock];
  const int i0 = lb * kBlock, i = i0 + threadIdx.x;
  double w[4], sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, mn[4], mx[4];
  double2 ci = make_double2(0.0, 0.0);
  if (i < N) {
    ld4(W, i, w);
    ci = cxy[i];
    st4(sW, threadIdx.x, w);
    sX[threadIdx.x] = ci;
  }
  __syncthreads();
  if (i >= N) return;
  const int s = i / kSlice, l = i % kSlice;
  const int base = soff[s], wid = swid[s];
  for (int v = 0; v < 4; ++v) mn[v] = mx[v] = w[v];
  // four neighbours per round: their ids, then their centres and states are
  // all in flight before the first one is used (memory-level parallelism for a
  // gather-latency-bound loop); the remainder goes one at a time
  auto accumulate = [&](unsigned code, const double2 cj, const double q[4]) {
    double xj = cj.x, yj = cj.y;
    const int sh = int(code >> 27);
    if (sh) {
      xj = xj + st.sx[sh - 1];
      yj = yj + st.sy[sh - 1];
    }
    lsq_term<LIM, PMODE>(xj, yj, ci, q, w, p, sx, sy, mn, mx);
  };
  int k = 0;
  for (; k + 4 <= wid; k += 4) {
    unsigned code[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) code[u] = sid[base + long(k + u) * kSlice + l];
    double2 cj[4];
    double q[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = int(code[u] & 0x07ffffffu);
      const unsigned lj = unsigned(j - i0);
      if (lj < unsigned(kBlock)) {
        cj[u] = sX[lj];
        ld4(sW, lj, q[u]);
      } else {
        cj[u] = cxy[j];
        ld4(W, j, q[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) accumulate(code[u], cj[u], q[u]);
  }
  for (; k < wid; ++k) {
    const unsigned code = sid[base + long(k) * kSlice + l];
    const int j = int(code & 0x07ffffffu);
    const unsigned lj = unsigned(j - i0);
    double q[4];
    double2 cj;
    if (lj < unsigned(kBlock)) {
      cj = sX[lj];
      ld4(sW, lj, q);
    } else {
      cj = cxy[j];
      ld4(W, j, q);
    }
    accumulate(code, cj, q);
  }
  double gx[4], gy[4];
  lsq_apply_ginv(ginv[i], ginv[N + i], ginv[2 * long(N) + i], sx, sy, gx, gy);
  double g8[8] = {gx[0], gx[1], gx[2], gx[3], gy[0], gy[1], gy[2], gy[3]};
  st8(grad, i, g8);
  if (LIM) {
    // batch mode with some instances frozen: find the cell's instance in the
    // offset table (ioff[0] = 0 <= i < ioff[ninst] = N) and keep its value
    bool keep = false;
    if (ioff) {
      int lo = 0, hi = ninst;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (ioff[mid] <= i) lo = mid;
        else hi = mid;
      }
      keep = frz[lo] != 0;
    }
    if (!keep) phi[i] = cell_limiter(lt, nvrt[i], N, i, loff, vol[i], w, mn, mx, gx, gy);
  }
}

// interior face flux with L = c1, R = c2 (shared by every flux kernel).
// The reconstruction (src/residual.f90:87-92) is written with explicit FMAs
// so that it rounds the same in every kernel it is inlined into, whatever
// the compiler's contraction choices there.
__device__ __forceinline__ void interior_flux_v(const Phys &P, const double g[8], const double wL[4],
                                                const double wR[4], const double gL[8], const double gR[8],
                                                double phL, double phR, double fa[4], double &wsa) {
  const double k2 = P.kappa / 2.0, k1 = 1.0 - P.kappa;
  double pL[4], pR[4];
  for (int v = 0; v < 4; ++v) {
    const double gC = wR[v] - wL[v];
    const double dL = fma(g[3], gL[v], g[4] * gL[4 + v]);
    const double dR = fma(g[5], gR[v], g[6] * gR[4 + v]);
    pL[v] = fma(phL, fma(k2, gC, k1 * dL), wL[v]);
    pR[v] = fma(phR, fma(-k2, gC, k1 * dR), wR[v]);
  }
  double fl[4], ws;
  roe_flux(P.gamma, pL, pR, g[0], g[1], fl, ws);
  const double A = g[2];
  for (int v = 0; v < 4; ++v) fa[v] = fl[v] * A;
  wsa = ws * A;
}

__device__ __forceinline__ void interior_flux(const Phys &P, int2 c, const double g[8], const double *__restrict__ W,
                                              const double *__restrict__ grad, const double *__restrict__ phi,
                                              double fa[4], double &wsa) {
  double wL[4], wR[4], gL[8], gR[8];
  ld4(W, c.x, wL);
  ld4(W, c.y, wR);
  ld8(grad, c.x, gL);
  ld8(grad, c.y, gR);
  interior_flux_v(P, g, wL, wR, gL, gR, phi[c.x], phi[c.y], fa, wsa);
}

__device__ __forceinline__ void boundary_flux_v(const Phys &P, double t, int ty, const double g[8], const double w[4],
                                                const double gr[8], double ph, double fa[4], double &wsa,
                                                const double *__restrict__ pinf) {
  double pL[4], pR[4];
  for (int v = 0; v < 4; ++v) pL[v] = fma(ph, fma(g[3], gr[v], g[4] * gr[4 + v]), w[v]);
  const double nx = g[0], ny = g[1];
  if (ty == FVS_BC_FREESTREAM) {
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
  for (int v = 0; v < 4; ++v) fa[v] = fl[v] * A;
  wsa = ws * A;
}

__device__ __forceinline__ void boundary_flux(const Phys &P, double t, int c, int ty, const double g[8],
                                              const double *__restrict__ W, const double *__restrict__ grad,
                                              const double *__restrict__ phi, double fa[4], double &wsa,
                                              const double *__restrict__ pinf) {
  double w[4], gr[8];
  ld4(W, c, w);
  ld8(grad, c, gr);
  boundary_flux_v(P, t, ty, g, w, gr, phi[c], fa, wsa, pinf);
}

// gather of the block's face fluxes from LDS in the reference's per-cell
// order (src/residual.f90:97-101,153-155), -R/vol (164-166), local time step
// (src/runge_kutta.f90:433), stage update, next primitive state, and the
// last stage's squared increments (per-block sums for the residual norm)
// fixed-shape tree over NT threads (deterministic); with NT = 2 kBlock the
// upper half contributes exact zeros, so the sums equal the kBlock tree's
template <int NT>
__device__ __forceinline__ void block_sum4_t(double v[4], double *red) {
  const int t = threadIdx.x;
  for (int k = 0; k < 4; ++k) red[k * NT + t] = v[k];
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < 4; ++k) red[k * NT + t] = red[k * NT + t] + red[k * NT + t + s];
    __syncthreads();
  }
  for (int k = 0; k < 4; ++k) v[k] = red[k * NT];
}

// The cell-side inputs of the stage update, loaded before the face fluxes
// are complete (their latency overlaps the barrier): the cell's gather
// entries (at most 4 faces), its volume, the step-start state and the stage
// state or accumulator the update reads, and its local time step.
struct UpdIn {
  int sv[4];
  double V, dl;
  double u0[4], x[4];  // x: SSP-RK3 stage state, else accumulator (0 on the first stage)
};

// Ux: where SSP-RK3 reads the stage state (the stage-state buffer the
// update writes in place).  x_given: the caller fills in the SSP-RK3 stage
// state (and on the first stage the step-start state) itself.
__device__ __forceinline__ void block_update_load(int N, int i, int gw, const Upd &A, const short *__restrict__ lgat,
                                                  const double *__restrict__ vol, const double *__restrict__ Ua,
                                                  const double *__restrict__ Ux, const double *__restrict__ Facc,
                                                  const double *__restrict__ dtl, UpdIn &in, bool x_given = false) {
#pragma unroll
  for (int k = 0; k < 4; ++k) in.sv[k] = (k < gw) ? lgat[long(k) * N + i] : 0;
  in.V = vol[i];
  in.dl = (A.steady && !A.first) ? dtl[i] : 0.0;
  // a stage of the two-register form with a = 0 after the first does not
  // read the step-start state (exact zeros keep a * u0 exact)
  const bool u0_unused = A.mode == M_SSP3 && A.a == 0.0 && !A.first;
  if (u0_unused) {
    for (int v = 0; v < 4; ++v) in.u0[v] = 0.0;
  } else if (!(x_given && A.first)) {
    ld4(Ua, i, in.u0);
  }
  if (A.mode == M_SSP3) {
    if (!x_given) {
      if (A.first) for (int v = 0; v < 4; ++v) in.x[v] = in.u0[v];
      else ld4(Ux, i, in.x);
    }
  } else {
    if (A.first) for (int v = 0; v < 4; ++v) in.x[v] = 0.0;
    else ld4(Facc, i, in.x);
  }
}

// The launch-block list of a batch with masked instances (nbl slots, the
// first gridDim.x of them the launched blocks: those holding a cell of an
// active instance) is followed by 8 words per block: bit l set when the
// block's cell l belongs to an active instance.
__device__ __forceinline__ const unsigned *active_mask(const int *blist, int nbl) {
  return reinterpret_cast<const unsigned *>(blist + nbl);
}

// NT threads per workgroup, CB cells per block (CB <= NT); instance masks
// exist for 256-cell blocks only
template <int NT, int CB = kBlock>
__device__ __forceinline__ void block_update_finish(int N, int lb, const Upd &A, const UpdIn &in,
                                                    const double *__restrict__ lF, const double *__restrict__ lS,
                                                    double *__restrict__ Ua, double *__restrict__ Ub,
                                                    double *__restrict__ Facc, double *__restrict__ Wout,
                                                    double *__restrict__ dtl, double *__restrict__ part,
                                                    double *__restrict__ d2out, double *__restrict__ red,
                                                    const unsigned *__restrict__ amask) {
  const int i = lb * CB + threadIdx.x;
  double d2[4] = {0, 0, 0, 0};
  // amask (batches with masked instances): bit l of the block's eight words
  // is set when its cell l belongs to an active instance; a masked cell
  // keeps its state (the update is computed, nothing is stored) and
  // contributes zero increments
  const bool on = !amask || threadIdx.x >= CB || ((amask[8 * lb + (threadIdx.x >> 5)] >> (threadIdx.x & 31)) & 1u);
  if (threadIdx.x < CB && i < N) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0}, wacc = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int s = in.sv[k];
      if (s == 0) break;
      const int f = (s > 0 ? s : -s) - 1;
      if (s > 0)
        for (int v = 0; v < 4; ++v) acc[v] = acc[v] + lF[4 * f + v];
      else
        for (int v = 0; v < 4; ++v) acc[v] = acc[v] - lF[4 * f + v];
      if (A.steady && A.first) wacc = wacc + lS[f];
    }
    const double V = in.V;
    double R[4];
    const double iV = 1.0 / V;  // one division per cell (src/residual.f90:164-166 divides each)
    for (int v = 0; v < 4; ++v) R[v] = -acc[v] * iV;
    double hs = A.h;
    if (A.steady) {
      double dl;
      if (A.first) {
        dl = A.cfl * V / (0.5 * wacc);
        if (on) dtl[i] = dl;
      } else {
        dl = in.dl;
      }
      hs = dl * A.h;
    }
    const double *u0 = in.u0;
    double un[4];
    if (A.mode == M_SSP3) {
      const double *uc = in.x;
      for (int v = 0; v < 4; ++v) un[v] = A.a * u0[v] + A.b * (uc[v] + hs * R[v]);
    } else if (A.mode == M_SSP42) {
      const double *fo = in.x;
      for (int v = 0; v < 4; ++v) un[v] = u0[v] + hs * (A.coef * R[v] + fo[v]);
      if (!A.last && on) {
        double fn[4];
        for (int v = 0; v < 4; ++v) fn[v] = fo[v] + R[v];
        st4(Facc, i, fn);
      }
    } else {
      const double *fo = in.x;
      double fn[4];
      for (int v = 0; v < 4; ++v) fn[v] = fo[v] + A.coef * R[v];
      if (A.use_acc) for (int v = 0; v < 4; ++v) un[v] = u0[v] + hs * fn[v];
      else for (int v = 0; v < 4; ++v) un[v] = u0[v] + hs * R[v];
      if (!A.last && on) st4(Facc, i, fn);
    }
    if (A.last) {
      if (on) {
        st4(Ua, i, un);
        for (int v = 0; v < 4; ++v) {
          const double d = fabs(un[v] - u0[v]);
          d2[v] = d * d;
        }
      }
      if (d2out) st4(d2out, i, d2);
    } else if (A.mode == M_SSP3 && on) {
      st4(Ub, i, un);
    }
    if (Wout && on) {
      double w[4];
      cons2prim(un, A.gamma, w);
      st4(Wout, i, w);
    }
  }
  if (A.last) {
    __syncthreads();  // red may alias the face fluxes just gathered
    block_sum4_t<NT>(d2, red);
    if (threadIdx.x == 0)
      for (int v = 0; v < 4; ++v) part[4 * lb + v] = d2[v];
  }
}

// fused face fluxes (LDS) + ordered gather + stage update; reads the stage's
// primitive state from Win and writes the next one to Wout (separate buffers:
// neighbouring blocks still read Win).  phi_one: no limiter, every limiter
// value is 1 and is not loaded.  With cons (SSP-RK3 only), Win is the
// stage's conservative input state and no primitive state is written (Wout
// null): the block's own cells are staged unconverted, every face side is
// converted as it is read (from LDS or HBM), the update takes its stage state
// from the staged copy and writes the new one to Ub (a buffer other than
// Win).  Converted copies would need another 8 KB of LDS or 8 registers per
// lane, a wavefront per SIMD less.
// BS cells (and threads) per workgroup; the per-block face lists (boff,
// lrec, lgeo, lgat) are those built for blocks of BS cells.
template <int BS>
__global__ __launch_bounds__(BS) void k_flux_update(int N, int nb, int gw, int maxf, Upd A, Phys P, double t,
                                                        const int *__restrict__ boff, const int2 *__restrict__ lrec,
                                                        const double *__restrict__ lgeo, long L,
                                                        const short *__restrict__ lgat,
                                                        const double *__restrict__ Win,
                                                        const double *__restrict__ grad,
                                                        const double *__restrict__ phi,
                                                        const double *__restrict__ vol, double *__restrict__ Ua,
                                                        double *__restrict__ Ub, double *__restrict__ Facc,
                                                        double *__restrict__ Wout, double *__restrict__ dtl,
                                                        double *__restrict__ part, const int *__restrict__ blist,
                                                        const double *__restrict__ pinf_inst,
                                                        double *__restrict__ d2out, int phi_one, int cons) {
  extern __shared__ double lds[];
  double *lF = lds;             // [maxf][4] flux*area (later the block reduction)
  double *lS = lds + 4 * maxf;  // [maxf]    wave speed*area, only when the stage sets local time steps
  const bool ws = A.steady && A.first;
  __shared__ double sG[8 * BS];  // gradients of the block's own cells
  __shared__ double sW[4 * BS];  // and their states (primitive; with cons conservative)
  const int lb = blist ? blist[blockIdx.x] : xcd_block(nb);
  const int nbl = nb < 0 ? -nb : nb;
  if (lb >= nbl) return;  // the whole workgroup leaves before any barrier
  const int i0 = lb * BS;
  const int f0 = boff[lb], nf = boff[lb + 1] - f0;
  // the first face record is in flight while the block's own cells are
  // staged, and so are its limiter values and its outside side's state and
  // gradient (a face touching the block has at most one cell outside it), so
  // that the first face waits on no global load after the barrier
  int f = threadIdx.x;
  int2 c = make_int2(0, 0);
  double g[8];
  g[7] = 0.0;
  double wO[4] = {0, 0, 0, 0}, gO[8] = {0, 0, 0, 0, 0, 0, 0, 0}, pO[2] = {1.0, 1.0};
  int2 c2 = make_int2(0, 0);  // the second face's cells, for its gathers' early start
  if (f + BS < nf) c2 = lrec[f0 + f + BS];
  if (f < nf) {
    c = lrec[f0 + f];
    for (int k = 0; k < 7; ++k) g[k] = lgeo[k * L + f0 + f];
    if (!phi_one) {
      pO[0] = phi[c.x];
      if (c.y >= 0) pO[1] = phi[c.y];
    }
    if (c.y >= 0) {
      const int co = unsigned(c.x - i0) >= unsigned(BS) ? c.x : (unsigned(c.y - i0) >= unsigned(BS) ? c.y : -1);
      if (co >= 0) {
        ld4(Win, co, wO);
        ld8(grad, co, gO);
      }
    }
  }
  if (i0 + int(threadIdx.x) < N) {
    double gg[8], ww[4];
    ld8(grad, i0 + threadIdx.x, gg);
    ld4(Win, i0 + threadIdx.x, ww);
    st8(sG, threadIdx.x, gg);
    st4(sW, threadIdx.x, ww);
  }
  __syncthreads();
  // one face's flux into LDS; pre: its outside side is the prefetched one
  auto face = [&](int fs, int2 cs, const double gs[8], bool pre) __attribute__((always_inline)) {
    double fa[4], wsa;
    if (cs.y >= 0) {
      double wL[4], wR[4], gL[8], gR[8];
      const unsigned la = unsigned(cs.x - i0), lc = unsigned(cs.y - i0);
      if (la < unsigned(BS)) {
        ld4(sW, la, wL);
        ld8(sG, la, gL);
      } else if (pre) {
        for (int v = 0; v < 4; ++v) wL[v] = wO[v];
        for (int v = 0; v < 8; ++v) gL[v] = gO[v];
      } else {
        ld4(Win, cs.x, wL);
        ld8(grad, cs.x, gL);
      }
      if (lc < unsigned(BS)) {
        ld4(sW, lc, wR);
        ld8(sG, lc, gR);
      } else if (pre) {
        for (int v = 0; v < 4; ++v) wR[v] = wO[v];
        for (int v = 0; v < 8; ++v) gR[v] = gO[v];
      } else {
        ld4(Win, cs.y, wR);
        ld8(grad, cs.y, gR);
      }
      make_prim(cons, A.gamma, wL);
      make_prim(cons, A.gamma, wR);
      const double phL = phi_one ? 1.0 : (pre ? pO[0] : phi[cs.x]);
      const double phR = phi_one ? 1.0 : (pre ? pO[1] : phi[cs.y]);
      interior_flux_v(P, gs, wL, wR, gL, gR, phL, phR, fa, wsa);
    } else {
      const int code = -1 - cs.y;
      double w[4], gr[8];
      ld4(sW, cs.x - i0, w);  // a boundary face's cell is one of the block's own
      ld8(sG, cs.x - i0, gr);
      make_prim(cons, A.gamma, w);
      boundary_flux_v(P, t, code & 3, gs, w, gr, phi_one ? 1.0 : (pre ? pO[0] : phi[cs.x]), fa, wsa,
                      pinf_inst ? pinf_inst + 4 * (code >> 2) : P.pinf);
    }
    for (int v = 0; v < 4; ++v) lF[4 * fs + v] = fa[v];
    if (ws) lS[fs] = wsa;
  };
  if (f < nf) face(f, c, g, true);
  for (f += BS; f < nf; f += BS) {
    c = (f == int(threadIdx.x) + BS) ? c2 : lrec[f0 + f];
    for (int k = 0; k < 7; ++k) g[k] = lgeo[k * L + f0 + f];
    face(f, c, g, false);
  }
  UpdIn in;
  const int ic = i0 + threadIdx.x;
  if (ic < N) {
    block_update_load(N, ic, gw, A, lgat, vol, Ua, Ub, Facc, dtl, in, cons != 0);
    if (cons) {  // the stage input staged in LDS (the step-start state too on the first stage)
      ld4(sW, threadIdx.x, in.x);
      if (A.first)
        for (int v = 0; v < 4; ++v) in.u0[v] = in.x[v];
    }
  }
  __syncthreads();
  block_update_finish<BS, BS>(N, lb, A, in, lF, lS, Ua, Ub, Facc, Wout, dtl, part, d2out, lF,
                              blist ? active_mask(blist, nbl) : nullptr);
}

// Whole RK stage in one kernel per 256-cell block: the block stages, in LDS,
// the state and centre of every cell its gradients read (C2: the block's
// cells, then the other cells of its faces -- together C1 -- then the rest
// of their stencils), computes the gradient (and limiter) of the C1 cells
// from LDS, the fluxes of its faces from LDS, and the gather + update of its
// own cells, with 512 threads per 256-cell block.  Halo gradients are
// recomputed by every block that needs them
// (about 19% more gradients on triangle meshes), so gradients never travel
// through HBM.  Arithmetic and summation order are those of k_grad_lsqc and
// k_flux_update, so the results are bit-identical to them.
//
// Per block b: C2 cells s2list[s2off[b] .. s2off[b+1]), the first s2n1[b]
// of them C1 (its own cells first, in order); C1 stencils in slices of 64
// cells (slice sl0[b] + q/64, width slw, entries slid[sleo + k*64 + q%64] =
// local C2 index << 4 | periodic-shift code); listed face f's cells as local
// indices blf[f] = c1 | c2 << 16 (c2 = 0xffff on a boundary face).
// phisrc: 0 -> limiter values 1 (no limiter), 1 -> read phi (frozen); with
// LIM the limiter is evaluated and written for the block's own cells.
// One stencil entry of an LDS-staged LSQ gradient: local code (index << 4 |
// periodic shift) of neighbour j of the cell at centre ci with state w.
// The term itself is the shared lsq_term.
template <int LIM, int PMODE>
__device__ __forceinline__ void lds_accum(unsigned code, const double2 ci, const double w[4],
                                          const double *__restrict__ lW, const double2 *__restrict__ lX,
                                          const ShiftTable &st, double p, double sx[4], double sy[4], double mn[4],
                                          double mx[4]) {
  const int j = int(code >> 4), sh = int(code & 15u);
  const double2 cj = lX[j];
  double q4[4];
  ld4(lW, j, q4);
  double xj = cj.x, yj = cj.y;
  if (sh) {
    xj = xj + st.sx[sh - 1];
    yj = yj + st.sy[sh - 1];
  }
  lsq_term<LIM, PMODE>(xj, yj, ci, q4, w, p, sx, sy, mn, mx);
}

// G^-1 product (G = its 11, 12, 22 entries) and limiter of a gathered
// cell.  Returns the limiter value: evaluated with LIM (unless the cell's
// instance is frozen; written to phi when own), else phi[id] when phisrc is
// 1, else 1.
template <int LIM>
__device__ __forceinline__ double grad_finish(int id, int N, const double w[4], const double sx[4],
                                              const double sy[4], const double mn[4], const double mx[4],
                                              const double G[3], double g8[8], int phisrc,
                                              double *__restrict__ phi, int lt, const int *__restrict__ nvrt,
                                              const double2 *__restrict__ loff, const double *__restrict__ vol,
                                              const int *__restrict__ ioff, const int *__restrict__ frz, int ninst,
                                              bool own) {
  double gx[4], gy[4];
  lsq_apply_ginv(G[0], G[1], G[2], sx, sy, gx, gy);
  for (int v = 0; v < 4; ++v) {
    g8[v] = gx[v];
    g8[4 + v] = gy[v];
  }
  bool keep = !LIM;
  if (LIM && ioff) {
    int lo = 0, hi = ninst;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (ioff[mid] <= id) lo = mid;
      else hi = mid;
    }
    keep = frz[lo] != 0;
  }
  if (!keep) {
    const double ph = cell_limiter(lt, nvrt[id], N, id, loff, vol[id], w, mn, mx, gx, gy);
    if (own) phi[id] = ph;
    return ph;
  }
  return phisrc == 1 ? phi[id] : 1.0;
}

// Gradient of local cell q (global id) from states and centres staged in
// LDS, with its stencil as local codes (index << 4 | periodic shift) in a
// 64-cell slice starting at sl (width wid).  Same arithmetic and order as
// k_grad_lsqc.
template <int LIM, int PMODE>
__device__ __forceinline__ double lds_gradient(int q, int id, int N, const double *__restrict__ lW,
                                               const double2 *__restrict__ lX, const unsigned short *__restrict__ sl,
                                               int wid, const ShiftTable &st, double p,
                                               const double *__restrict__ ginv, double g8[8], int phisrc,
                                               double *__restrict__ phi, int lt, const int *__restrict__ nvrt,
                                               const double2 *__restrict__ loff, const double *__restrict__ vol,
                                               const int *__restrict__ ioff, const int *__restrict__ frz, int ninst,
                                               bool own) {
  const int l = q % kSlice;
  const double2 ci = lX[q];
  double w[4], sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, mn[4], mx[4];
  ld4(lW, q, w);
  for (int v = 0; v < 4; ++v) mn[v] = mx[v] = w[v];
  for (int k = 0; k < wid; ++k)
    lds_accum<LIM, PMODE>(sl[long(k) * kSlice + l], ci, w, lW, lX, st, p, sx, sy, mn, mx);
  const double G[3] = {ginv[id], ginv[N + id], ginv[2 * long(N) + id]};
  return grad_finish<LIM>(id, N, w, sx, sy, mn, mx, G, g8, phisrc, phi, lt, nvrt, loff, vol, ioff, frz, ninst,
                          own);
}

// LSQ gradient (and limiter) of the block's own cells with the cells its
// stencils read staged in LDS first: the block's own cells (coalesced) at
// local 0..255, the halo (the outside cells its stencils read, one gather
// each) from 256 on.  The stencils are 16-bit local codes, two per 32-bit
// word (entries 2kk and 2kk+1 of lane l of slice s at lcode[poff[s] + 64 kk
// + l]); a thread loads the first kGradRegs of its own while the block is
// staged, so the LDS loop waits on no global load.  Same arithmetic and order
// as k_grad_lsqc, so the gradients are bit-identical to it.  With cons, W
// holds conservative states, converted as they are staged.
constexpr int kGradRegs = 16;
constexpr int kHaloRounds = 4;  // halo list entries per thread loaded up front
// BS cells (and threads) per block, with the lists built for blocks of BS
// cells (without the limiter the kernel fits 72 registers: seven wavefronts
// per SIMD)
template <int LIM, int PMODE, int BS = kBlock>
__global__ __launch_bounds__(BS, LIM ? 1 : 7) void k_grad_lsqh(int N, int nb, int nl, const double *__restrict__ W,
                                                      int cons, double gamma, const int *__restrict__ ghpad,
                                                      const int *__restrict__ poff, const int *__restrict__ swid,
                                                      const unsigned *__restrict__ lcode,
                                                      const double2 *__restrict__ cxy,
                                                      const double *__restrict__ ginv, ShiftTable st, double p,
                                                      double *__restrict__ grad, double *__restrict__ phi, int lt,
                                                      const int *__restrict__ nvrt, const double2 *__restrict__ loff,
                                                      const double *__restrict__ vol, const int *__restrict__ blist,
                                                      const int *__restrict__ ioff, const int *__restrict__ frz,
                                                      int ninst) {
  extern __shared__ double lds[];
  double *lW = lds;                                          // [nl][4]
  double2 *lX = reinterpret_cast<double2 *>(lds + 4 * nl);   // [nl]
  const int lb = blist ? blist[blockIdx.x] : xcd_block(nb);
  if (lb >= (nb < 0 ? -nb : nb)) return;  // the whole workgroup leaves before the barrier
  const int i0 = lb * BS, id = i0 + threadIdx.x;
  const bool act = id < N;
  const int s = id / kSlice, wid = act ? swid[s] : 0;
  const unsigned *lc = lcode + (act ? poff[s] : 0) + threadIdx.x % kSlice;
  unsigned cr[kGradRegs / 2];
#pragma unroll
  for (int kk = 0; kk < kGradRegs / 2; ++kk) cr[kk] = (2 * kk < wid) ? lc[kk * kSlice] : 0u;
  const int *hl = ghpad + long(lb) * (nl - BS);  // the block's halo list, -1 after its end
  double G[3] = {0.0, 0.0, 0.0};
  if (act) {
    G[0] = ginv[id];
    G[1] = ginv[N + id];
    G[2] = ginv[2 * long(N) + id];
    const double2 xy = cxy[id];
    double w[4];
    ld4(W, id, w);
    lX[threadIdx.x] = xy;
    make_prim(cons, gamma, w);
    st4(lW, threadIdx.x, w);
  }
  // the thread's first kHaloRounds list entries load together (no round
  // waits on the previous one to learn whether the list goes on); longer
  // halos continue entry by entry
  const int maxh = nl - BS;
  int jh[kHaloRounds];
#pragma unroll
  for (int r = 0; r < kHaloRounds; ++r) {
    const int q = int(threadIdx.x) + r * BS;
    jh[r] = q < maxh ? hl[q] : -1;
  }
#pragma unroll
  for (int r = 0; r < kHaloRounds; ++r) {
    const int q = int(threadIdx.x) + r * BS, j = jh[r];
    if (j < 0) break;
    const double2 xy = cxy[j];
    double w[4];
    ld4(W, j, w);
    lX[BS + q] = xy;
    make_prim(cons, gamma, w);
    st4(lW, BS + q, w);
  }
  for (int q = int(threadIdx.x) + kHaloRounds * BS; q < maxh && jh[kHaloRounds - 1] >= 0; q += BS) {
    const int j = hl[q];
    if (j < 0) break;
    const double2 xy = cxy[j];
    double w[4];
    ld4(W, j, w);
    lX[BS + q] = xy;
    make_prim(cons, gamma, w);
    st4(lW, BS + q, w);
  }
  __syncthreads();
  if (!act) return;
  const double2 ci = lX[threadIdx.x];
  double w[4], sx[4] = {0, 0, 0, 0}, sy[4] = {0, 0, 0, 0}, mn[4], mx[4];
  ld4(lW, threadIdx.x, w);
  for (int v = 0; v < 4; ++v) mn[v] = mx[v] = w[v];
#pragma unroll
  for (int k = 0; k < kGradRegs; ++k)
    if (k < wid) lds_accum<LIM, PMODE>((cr[k >> 1] >> (16 * (k & 1))) & 0xffffu, ci, w, lW, lX, st, p, sx, sy, mn, mx);
  for (int k = kGradRegs; k < wid; ++k)
    lds_accum<LIM, PMODE>((lc[(k >> 1) * kSlice] >> (16 * (k & 1))) & 0xffffu, ci, w, lW, lX, st, p, sx, sy, mn, mx);
  double g8[8];
  grad_finish<LIM>(id, N, w, sx, sy, mn, mx, G, g8, 0, phi, lt, nvrt, loff, vol, ioff, frz, ninst, true);
  st8(grad, id, g8);
}

constexpr int kStageThreads = 2 * kBlock;  // two threads per cell: each phase fits one round

// Everything k_stage reads besides the stage's own update constants and its
// input / output primitive states (one struct, so the same block body serves
// the per-stage launch and the persistent whole-step kernel).
struct StageArgs {
  int N, nb, gw, maxf, max1, max2;
  Phys P;
  double t;
  const int *boff, *bfaces;
  const unsigned *blf;
  const short *lgat;
  int Fi;
  const double *fgeo;
  const int *btype;
  const double *bgeo;
  const int *s2off, *s2n1, *s2list, *sl0, *slw;
  const long *sleo;
  const unsigned short *slid;
  const double2 *cxy;
  const double *ginv;
  ShiftTable st;
  double p;
  int phisrc;
  double *phi;
  int lt;
  const int *nvrt;
  const double2 *loff;
  const double *vol;
  const int *ioff, *frz;
  int ninst;
  double *Ua, *Ub, *Facc, *dtl, *part;
  const int *blist, *binst;
  const double *pinf_inst;
};

// One 256-cell block lb of a whole RK stage (see above).
template <int LIM, int PMODE>
__device__ __forceinline__ void stage_block(const StageArgs &a, int lb, const Upd &A, const double *__restrict__ Win,
                                            double *__restrict__ Wout, double *__restrict__ d2out) {
  extern __shared__ double lds[];
  const int N = a.N, max1 = a.max1, max2 = a.max2, maxf = a.maxf;
  double *lW = lds;                                   // [max2][4]
  double2 *lX = reinterpret_cast<double2 *>(lds + 4 * max2);  // [max2]
  double *lG = lds + 6 * max2;                        // [max1][8]
  double *lP = lG + 8 * max1;                         // [max1]
  double *lF = lP + max1;                             // [maxf][4]
  double *lS = lF + 4 * maxf;                         // [maxf]
  const int nbl = a.nb < 0 ? -a.nb : a.nb;
  const int o2 = a.s2off[lb], n2 = a.s2off[lb + 1] - o2, n1 = a.s2n1[lb];
  const int i0 = lb * kBlock, nown = (N - i0 < kBlock) ? N - i0 : kBlock;
  // stage the state and centres of C2
  for (int q = threadIdx.x; q < n2; q += kStageThreads) {
    const int id = a.s2list[o2 + q];
    double w[4];
    ld4(Win, id, w);
    st4(lW, q, w);
    lX[q] = a.cxy[id];
  }
  __syncthreads();
  // gradients (and limiter values) of C1 from LDS
  for (int q = threadIdx.x; q < n1; q += kStageThreads) {
    const int id = a.s2list[o2 + q];
    const int si = a.sl0[lb] + q / kSlice;
    double g8[8];
    const double ph = lds_gradient<LIM, PMODE>(q, id, N, lW, lX, a.slid + a.sleo[si], a.slw[si], a.st, a.p, a.ginv,
                                               g8, a.phisrc, a.phi, a.lt, a.nvrt, a.loff, a.vol, a.ioff, a.frz,
                                               a.ninst, q < nown);
    st8(lG, q, g8);
    lP[q] = ph;
  }
  __syncthreads();
  // face fluxes from LDS
  const int f0 = a.boff[lb], nf = a.boff[lb + 1] - f0;
  for (int f = threadIdx.x; f < nf; f += kStageThreads) {
    const int slot = a.bfaces[f0 + f];
    const unsigned pr = a.blf[f0 + f];
    const int la = int(pr & 0xffffu), lc = int(pr >> 16);
    double fa[4], wsa, g[8];
    if (slot < a.Fi) {
      ld8(a.fgeo, slot, g);
      interior_flux(a.P, make_int2(la, lc), g, lW, lG, lP, fa, wsa);
    } else {
      const int b = slot - a.Fi;
      ld8(a.bgeo, b, g);
      boundary_flux(a.P, a.t, la, a.btype[b], g, lW, lG, lP, fa, wsa,
                    a.binst ? a.pinf_inst + 4 * a.binst[b] : a.P.pinf);
    }
    for (int v = 0; v < 4; ++v) lF[4 * f + v] = fa[v];
    lS[f] = wsa;
  }
  __syncthreads();
  // the staged states are no longer read: their LDS holds the reduction
  {
    UpdIn in;
    const int ic = lb * kBlock + threadIdx.x;
    if (threadIdx.x < kBlock && ic < N)
      block_update_load(N, ic, a.gw, A, a.lgat, a.vol, a.Ua, a.Ub, a.Facc, a.dtl, in);
    block_update_finish<kStageThreads>(N, lb, A, in, lF, lS, a.Ua, a.Ub, a.Facc, Wout, a.dtl, a.part, d2out, lds,
                                       a.blist ? active_mask(a.blist, nbl) : nullptr);
  }
}

template <int LIM, int PMODE>
__global__ __launch_bounds__(kStageThreads) void k_stage(StageArgs a, Upd A, const double *__restrict__ Win,
                                                         double *__restrict__ Wout, double *__restrict__ d2out) {
  const int lb = a.blist ? a.blist[blockIdx.x] : xcd_block(a.nb);
  const int nbl = a.nb < 0 ? -a.nb : a.nb;
  if (lb >= nbl) return;  // the whole workgroup leaves before any barrier
  stage_block<LIM, PMODE>(a, lb, A, Win, Wout, d2out);
}

// Thread t's share of an instance's squared increments: cells a + t, a + t
// + kBlock, ... < b added in that order, with kInstLoads loads in flight
// before their adds (an instance holds up to 2^18 cells, one workgroup walks
// it, and the loads' latency would otherwise be paid once per cell)
constexpr int kInstLoads = 8;
__device__ __forceinline__ void inst_share(int a, int b, int t, const double *__restrict__ d2, double s[4]) {
  for (int i0 = a + t; i0 < b; i0 += kInstLoads * kBlock) {
    double q[kInstLoads][4];
#pragma unroll
    for (int u = 0; u < kInstLoads; ++u) {
      const int i = i0 + u * kBlock;
      if (i < b) ld4(d2, i, q[u]);
    }
#pragma unroll
    for (int u = 0; u < kInstLoads; ++u)
      if (i0 + u * kBlock < b)
        for (int v = 0; v < 4; ++v) s[v] = s[v] + q[u][v];
  }
}

// per-instance residual norms of one step from the per-cell squared
// increments: one workgroup per instance, fixed-shape tree (deterministic)
__global__ __launch_bounds__(kBlock) void k_inst_norms(int ninst, const int *__restrict__ off,
                                                       const double *__restrict__ d2, const int *__restrict__ ctr,
                                                       double *__restrict__ out) {
  __shared__ double red[4 * kBlock];
  const int k = blockIdx.x;
  if (k >= ninst) return;
  const int a = off[k], b = off[k + 1];
  double s[4] = {0, 0, 0, 0};
  inst_share(a, b, threadIdx.x, d2, s);
  block_sum4(s, red);
  if (threadIdx.x == 0) {
    double *o = out + 4 * (long(ctr ? ctr[0] : 0) * ninst + k);
    for (int v = 0; v < 4; ++v) o[v] = sqrt(s[v] / double(b - a));
  }
}

__global__ void k_counter(int *__restrict__ ctr, int value, int add) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctr[0] = add ? ctr[0] + value : value;
}

// C_L / C_D per instance: thread k walks instance k's wall faces in order
__global__ void k_wall_forces_inst(int ninst, const int *__restrict__ woff, const int *__restrict__ wf,
                                   const int *__restrict__ bcell, const double *__restrict__ geo,
                                   const double *__restrict__ W, const double *__restrict__ grad,
                                   const double *__restrict__ mach_aoa, const double *__restrict__ pinf_inst,
                                   double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ninst) return;
  const double M = mach_aoa[2 * k], al = mach_aoa[2 * k + 1] * (kPiD / 180.0);
  const double pinf = pinf_inst[4 * k + 3];
  double cn = 0, ca = 0, cn1 = 0, ca1 = 0;
  for (int q = woff[k]; q < woff[k + 1]; ++q) {
    const int b = wf[q], c = bcell[b];
    double g[8], gr[8], w[4];
    ld8(geo, b, g);
    ld8(grad, c, gr);
    ld4(W, c, w);
    const double pw = w[3] + g[3] * gr[3] + g[4] * gr[7];
    const double cp = 2.0 / (M * M) * (pw - pinf);
    const double cp1 = 2.0 / (M * M) * (w[3] - pinf);
    cn = cn + cp * g[1] * g[2];
    ca = ca + cp * g[0] * g[2];
    cn1 = cn1 + cp1 * g[1] * g[2];
    ca1 = ca1 + cp1 * g[0] * g[2];
  }
  const double c = cos(al), sn = sin(al);
  out[4 * k] = cn * c - ca * sn;
  out[4 * k + 1] = cn * sn + ca * c;
  out[4 * k + 2] = cn1 * c - ca1 * sn;
  out[4 * k + 3] = cn1 * sn + ca1 * c;
}

// ---------------------------------------------------------------------------
// Persistent whole-step kernel: one cooperative launch runs nsteps time steps
// of the whole-stage block body over the launch-block list (the ragged batch's
// active blocks, or every block), with a grid-wide barrier after each stage
// and the step's residual norms reduced in the same launch.
// ---------------------------------------------------------------------------

// Grid-wide barrier of a co-resident launch: thread 0 of each workgroup
// counts itself in bar[0]; the last one resets the count and advances the
// generation bar[1]; the others poll the generation.  Each poll is bounded
// (spin polls, or as soon as another workgroup has given up and set *err):
// a workgroup that gives up records it in *err and returns false, so every
// wave of the launch ends.
__device__ __forceinline__ bool grid_barrier(unsigned *bar, unsigned nwg, unsigned &gen, unsigned spin, int *err) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    ok = 1;
    __threadfence();
    const unsigned arrived = atomicAdd(&bar[0], 1u);
    if (arrived == nwg - 1) {
      atomicExch(&bar[0], 0u);
      __threadfence();
      atomicAdd(&bar[1], 1u);
    } else {
      unsigned n = 0;
      while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
        __builtin_amdgcn_s_sleep(2);
        if (++n >= spin || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          atomicExch(err, 1);
          ok = 0;
          break;
        }
      }
    }
    gen = gen + 1;
    __threadfence();
  }
  __syncthreads();
  return ok != 0;
}

// The whole-mesh residual norms of one step (k_norm_final's summation:
// kNormThreads partial sums over strided partials, then its tree), computed
// by NT threads that each play kNormThreads / NT of its threads, so the bits
// are those of k_norm_final.  red: 4 kNormThreads doubles of LDS.
template <int NT>
__device__ __forceinline__ void norm_sum_block(int nparts, int N, const double *__restrict__ part,
                                               double *__restrict__ out, double *red) {
  constexpr int R = kNormThreads / NT;
  for (int r = 0; r < R; ++r) {
    const int vt = threadIdx.x + r * NT;
    double s[4] = {0, 0, 0, 0};
    for (int p0 = vt; p0 < nparts; p0 += kNormLoads * kNormThreads) {
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
    for (int v = 0; v < 4; ++v) red[v * kNormThreads + vt] = s[v];
  }
  __syncthreads();
  for (int h = kNormThreads / 2; h > 0; h >>= 1) {
    for (int r = 0; r < R; ++r) {
      const int vt = threadIdx.x + r * NT;
      if (vt < h)
        for (int v = 0; v < 4; ++v)
          red[v * kNormThreads + vt] = red[v * kNormThreads + vt] + red[v * kNormThreads + vt + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int v = 0; v < 4; ++v) out[v] = sqrt(red[v * kNormThreads] / double(N));
  __syncthreads();
}

// Instance k's residual norms of one step with k_inst_norms' summation (the
// first kBlock threads of the workgroup play its threads).  red: 4 kBlock
// doubles of LDS.
__device__ __forceinline__ void inst_norm_block(int k, const int *__restrict__ off, const double *__restrict__ d2,
                                                double *__restrict__ out, double *red) {
  const int t = threadIdx.x;
  const int a = off[k], b = off[k + 1];
  if (t < kBlock) {
    double s[4] = {0, 0, 0, 0};
    inst_share(a, b, t, d2, s);
    for (int v = 0; v < 4; ++v) red[v * kBlock + t] = s[v];
  }
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if (t < h)
      for (int v = 0; v < 4; ++v) red[v * kBlock + t] = red[v * kBlock + t] + red[v * kBlock + t + h];
    __syncthreads();
  }
  if (t == 0)
    for (int v = 0; v < 4; ++v) out[v] = sqrt(red[v * kBlock] / double(b - a));
  __syncthreads();
}

struct PersistCtl {
  Upd A[4];                 // the stages of one step
  int nst, nsteps, nlaunch;  // stages per step, steps, launch blocks
  double *W0, *W1;          // primitive states: stage k of the launch reads W0 when k is even
  double *d2buf;            // per-cell squared increments of a step's last stage
  double *dn;               // whole-mesh norms, one row per step (NULL: none)
  int nparts;
  double *dinst;            // per-instance norms (NULL: none)
  int ninst;
  const int *inst_off;
  int *ctr;                 // output row of the first step; advanced by nsteps at the end
  unsigned *bar;
  int *err;
  unsigned spin;
};

template <int LIM, int PMODE>
__global__ __launch_bounds__(kStageThreads) void k_steps_persistent(StageArgs a, PersistCtl c) {
  extern __shared__ double lds[];
  unsigned gen = 0;
  if (threadIdx.x == 0) gen = __hip_atomic_load(&c.bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int row0 = c.ctr[0];  // block 0 writes the counter only after the last barrier
  int par = 0;
  bool ok = true;
  for (int s = 0; s < c.nsteps && ok; ++s) {
    for (int k = 0; k < c.nst && ok; ++k) {
      const Upd A = c.A[k];
      const double *Win = par ? c.W1 : c.W0;
      double *Wout = par ? c.W0 : c.W1;
      for (int q = blockIdx.x; q < c.nlaunch; q += gridDim.x) {
        const int lb = a.blist ? a.blist[q] : q;
        stage_block<LIM, PMODE>(a, lb, A, Win, Wout, A.last ? c.d2buf : nullptr);
        __syncthreads();  // the next block's staging overwrites this one's LDS
      }
      ok = grid_barrier(c.bar, gridDim.x, gen, c.spin, c.err);
      par ^= 1;
    }
    if (!ok) break;
    // this step's norms; the partials and increments they read are next
    // written by the last stage of the next step, two or more barriers later
    const long row = long(row0) + s;
    if (c.dn && blockIdx.x == 0) norm_sum_block<kStageThreads>(c.nparts, a.N, a.part, c.dn + 4 * row, lds);
    if (c.dinst)
      for (int k = blockIdx.x; k < c.ninst; k += gridDim.x)
        inst_norm_block(k, c.inst_off, c.d2buf, c.dinst + 4 * (row * c.ninst + k), lds);
  }
  if (ok) ok = grid_barrier(c.bar, gridDim.x, gen, c.spin, c.err);
  if (ok && blockIdx.x == 0 && threadIdx.x == 0) c.ctr[0] = row0 + c.nsteps;
}

__global__ void k_barrier_reset(unsigned *bar, int *err) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    bar[0] = 0u;
    err[0] = 0;
  }
}
