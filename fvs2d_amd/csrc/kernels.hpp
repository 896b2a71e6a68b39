// Device kernels of the fvs2d residual / time-step path (gfx950, FP64).
//
// Data layout in HBM (cells in device order, see fvs_api.hip):
//   U0, U, F, W  double[N][4]  conservative state (step start / stage),
//                              RK accumulator, primitive state (rho,u,v,p)
//   grad         double[N][8]  d/dx of the 4 primitives, then d/dy
//   phi, dtl     double[N]
//   interior faces: int2 cells[Fi] (c1,c2), double geo[Fi][8] =
//                   (nx, ny, area, xf-xc1, yf-yc1, xf-xc2, yf-yc2, 0)
//   boundary faces: int cell[Fb], int type[Fb], double geo[Fb][8] =
//                   (nx, ny, area, xf-xc, yf-yc, xf, yf, 0)
//   FA[Fi+Fb][4], WSA[Fi+Fb]   flux*area and (wave speed)*area per face
//   gather table int gat[GW][N]: signed (face slot + 1), in the reference's
//                   per-cell summation order, 0 = unused
//   stencils in sliced-ELL form: slices of 64 cells (one wavefront), entry
//                   k of lane l at off[slice] + 64*k + l
#pragma once
#include <hip/hip_runtime.h>

namespace fvs {

constexpr double kPiD = 3.14159265358979323846;

struct Phys {
  double gamma, kappa;
  double pinf[4];
  double vx0, vy0, vK, vinf[4];
  int lvortex;
  int limiter;
};

// ---------------------------------------------------------------------------
// pointwise physics
// ---------------------------------------------------------------------------
// cvar -> pvar: src/data_solution.f90:72-86.  No multiply-add contraction:
// every kernel that converts a state (the update that writes it, or a
// gradient / flux kernel staging a conservative state) gets the same bits.
__device__ __forceinline__ void cons2prim(const double u[4], double g, double w[4]) {
#pragma clang fp contract(off)
  w[0] = u[0];
  w[1] = u[1] / u[0];
  w[2] = u[2] / u[0];
  w[3] = (g - 1.0) * (u[3] - 0.5 * w[0] * (w[1] * w[1] + w[2] * w[2]));
}

// Roe flux with Harten's entropy fix: src/flux_invscid.f90:37-136.
// The flux is 0.5*(F_L + F_R - sum_j |lambda_j| alpha_j r_j) with the four
// waves of the Roe-averaged state; returns 0.5*(|u_n| + a) in ws.  Quotients
// that share a divisor (1/rho_L, 1/(gamma-1), the Roe-average weight
// 1/(1+R), 1/a^2, Harten's 1/d) multiply by its reciprocal: each differs from the
// reference's division by at most an ulp, and an FP64 division costs about
// ten instructions, which this flux (about 1.7 per cell and stage) is
// bound by as much as by its loads.
__device__ __forceinline__ void roe_flux(double g, const double pL[4], const double pR[4], double nx, double ny,
                                         double f[4], double &ws) {
  const double tx = -ny, ty = nx;
  const double rL = pL[0], uL = pL[1], vL = pL[2], PL = pL[3];
  const double rR = pR[0], uR = pR[1], vR = pR[2], PR = pR[3];
  const double unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
  const double utL = uL * tx + vL * ty, utR = uR * tx + vR * ty;
  const double irL = 1.0 / rL, gi = 1.0 / (g - 1.0);
  const double aL = sqrt(g * PL * irL), aR = sqrt(g * PR / rR);
  const double kL = 0.5 * (uL * uL + vL * vL), kR = 0.5 * (uR * uR + vR * vR);
  const double HL = aL * aL * gi + kL, HR = aR * aR * gi + kR;
  const double RT = sqrt(rR * irL);
  const double rho = RT * rL;
  const double iR = 1.0 / (1.0 + RT);
  const double u = (uL + RT * uR) * iR;
  const double v = (vL + RT * vR) * iR;
  const double H = (HL + RT * HR) * iR;
  const double a = sqrt((g - 1.0) * (H - 0.5 * (u * u + v * v)));
  const double un = u * nx + v * ny, ut = u * tx + v * ty;
  const double drho = rR - rL, dp = PR - PL, dun = unR - unL, dut = utR - utL;
  const double ia2 = 1.0 / (a * a);
  double L0 = 0.5 * (dp - rho * a * dun) * ia2;
  double L1 = rho * dut;
  double L2 = drho - dp * ia2;
  double L3 = 0.5 * (dp + rho * a * dun) * ia2;
  double w0 = fabs(un - a), w1 = fabs(un), w3 = fabs(un + a);
  const double d = 1.0 / 5.0, id = 5.0;  // Harten's width and its reciprocal
  if (w0 < d) w0 = 0.5 * (w0 * w0 * id + d);
  if (w3 < d) w3 = 0.5 * (w3 * w3 * id + d);
  const double tke = 0.5 * (u * u + v * v);
  // wave amplitude * |speed|, then dissipation = sum_j s_j r_j
  const double s0 = w0 * L0, s1 = w1 * L1, s2 = w1 * L2, s3 = w3 * L3;
  // r_1 has no density component, so D0 takes three waves
  double D0 = s0 + s2 + s3, D1 = 0.0, D2 = 0.0, D3 = 0.0;
  D1 = D1 + s0 * (u - a * nx); D1 = D1 + s1 * tx; D1 = D1 + s2 * u; D1 = D1 + s3 * (u + a * nx);
  D2 = D2 + s0 * (v - a * ny); D2 = D2 + s1 * ty; D2 = D2 + s2 * v; D2 = D2 + s3 * (v + a * ny);
  D3 = D3 + s0 * (H - un * a); D3 = D3 + s1 * ut; D3 = D3 + s2 * tke; D3 = D3 + s3 * (H + un * a);
  const double mL = rL * unL, mR = rR * unR;
  const double fL1 = mL * uL + PL * nx, fL2 = mL * vL + PL * ny, fL3 = mL * HL;
  const double fR1 = mR * uR + PR * nx, fR2 = mR * vR + PR * ny, fR3 = mR * HR;
  f[0] = 0.5 * (mL + mR - D0);
  f[1] = 0.5 * (fL1 + fR1 - D1);
  f[2] = 0.5 * (fL2 + fR2 - D2);
  f[3] = 0.5 * (fL3 + fR3 - D3);
  ws = 0.5 * (fabs(un) + a);
}

// isentropic vortex: src/mms.f90:219-265
__device__ __forceinline__ void vortex_state(const Phys &P, double t, double x, double y, double pv[4]) {
  const double g = P.gamma, K = P.vK;
  const double Tinf = P.vinf[3] / P.vinf[0];
  const double dx = x - (P.vx0 + P.vinf[1] * t), dy = y - (P.vy0 + P.vinf[2] * t);
  const double r = sqrt(dx * dx + dy * dy);
  const double c = K / (2.0 * kPiD);
  const double e = exp(0.5 * (1.0 - r * r));
  pv[1] = P.vinf[1] - c * dy * e;
  pv[2] = P.vinf[2] + c * dx * e;
  const double temp = Tinf - (c * c) * (g - 1.0) / (2.0 * g) * exp(1.0 - r * r);
  pv[0] = pow(temp, 1.0 / (g - 1.0));
  pv[3] = pow(pv[0], g);
}

// manufactured solution (primitive state only): src/mms.f90:79-101,187-213
__device__ __forceinline__ void mms_state(double x, double y, double pv[4]) {
  pv[0] = 1.12 + 0.15 * sin((3.12 * kPiD) * x + (2.92 * kPiD) * y);
  pv[1] = 1.32 + 0.06 * sin((2.09 * kPiD) * x + (3.12 * kPiD) * y);
  pv[2] = 1.18 + 0.03 * sin((2.15 * kPiD) * x + (3.32 * kPiD) * y);
  pv[3] = 1.62 + 0.31 * sin((3.79 * kPiD) * x + (2.98 * kPiD) * y);
}

// Venkatakrishnan's eps^2 = (K h)^3 with K = 5 and h the diameter of the
// circle of the cell's area (src/gradient_limiter.f90:114-116)
__device__ __forceinline__ double venkat_eps2(double vol) {
  const double h = 2.0 * sqrt(vol / kPiD);
  const double t = 5.0 * h;
  return t * t * t;
}

// limiter function: src/gradient_limiter.f90:103-134
__device__ __forceinline__ double limiter_fn(int lt, double a, double b, double vol) {
  if (lt == 1) {
    const double eps2 = venkat_eps2(vol);
    return ((a * a + eps2) + 2.0 * b * a) / (a * a + 2.0 * (b * b) + a * b + eps2);
  } else if (lt == 2) {
    const double r = a / b;
    return (1.0 < r) ? 1.0 : r;
  } else {
    const double h = 2.0 * sqrt(vol / kPiD);
    const double t = 0.3 * h;
    const double eps2 = t * t * t;
    const double l = ((b * b + eps2) * a + (a * a + eps2) * b) / (a * a + b * b + 2.0 * eps2);
    return l / (b + eps2);
  }
}

}  // namespace fvs
