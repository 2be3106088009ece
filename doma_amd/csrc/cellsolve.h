// cellsolve.h -- per-cell discrete-ordinates solves (device functions).
//
// Cell solves: TWDiamondDifference3D.h:15-92, DiamondDifference3D.h:14-49,
// TWDiamondDifference2D.h:15-76, DiamondDifference2D.h:14-45,
// DiamondDifference1D.h:14-46.  The "exact" variants keep the reference's
// expression order with IEEE divides and no contraction; the "fast" variants
// hoist the weight reciprocals (1/a = d/(d-n) with the same clamp) and use FMA.
#pragma once

#include <hip/hip_runtime.h>

#include "doma_c.h"

namespace doma
{
// ------------------------------------------------------------------------- //
// Cell solves                                                                //
// ------------------------------------------------------------------------- //

template <typename R>
__device__ __forceinline__ R
twdd3d_exact(R q, R sig, R lx, R ly, R lz, R amu, R aeta, R axi, R & x, R & y, R & z)
{
#pragma clang fp contract(off)
  const R theta = R(0.9), theta_s = R(0.9);
  const R A = ly * lz;
  const R B = lx * lz;
  const R C = lx * ly;
  const R V = lx * ly * lz;
  R a = R(0.5);
  if (x > R(1e-8))
  {
    a = q * V * theta_s + (aeta * B * y + axi * C * z) * theta + amu * A * x;
    a /= (sig * V + R(2.0) * aeta * B + R(2.0) * axi * C) * x;
    a = R(-1.0) * (a - R(1.0));
  }
  R b = R(0.5);
  if (y > R(1e-8))
  {
    b = q * V * theta_s + (amu * A * x + axi * C * z) * theta + aeta * B * y;
    b /= (sig * V + R(2.0) * amu * A + R(2.0) * axi * C) * y;
    b = R(-1.0) * (b - R(1.0));
  }
  R c = R(0.5);
  if (z > R(1e-8))
  {
    c = q * V * theta_s + (amu * A * x + aeta * B * y) * theta + axi * C * z;
    c /= (sig * V + R(2.0) * amu * A + R(2.0) * aeta * B) * z;
    c = R(-1.0) * (c - R(1.0));
  }
  a = a < R(0.5) ? R(0.5) : a;
  a = a > R(1.0) ? R(1.0) : a;
  b = b < R(0.5) ? R(0.5) : b;
  b = b > R(1.0) ? R(1.0) : b;
  c = c < R(0.5) ? R(0.5) : c;
  c = c > R(1.0) ? R(1.0) : c;
  R psi = q * V + (amu * A / a * x) + (aeta * B / b * y) + (axi * C / c * z);
  psi /= sig * V + (amu * A / a) + (aeta * B / b) + (axi * C / c);
  x = (psi / a) - (((R(1.0) - a) / a) * x);
  y = (psi / b) - (((R(1.0) - b) / b) * y);
  z = (psi / c) - (((R(1.0) - c) / c) * z);
  return psi;
}

// Reciprocal of a normal, finite d: the hardware estimate r refined by one
// third-order step r (1 + e + e^2) with e = 1 - d r, whose relative error is
// e^3 (the same order of refinement as two Newton steps, in three dependent
// FMAs instead of four).
__device__ __forceinline__ double
fast_rcp(double d)
{
  const double r = __builtin_amdgcn_rcp(d);
  const double e = fma(-d, r, 1.0);
  return fma(r, fma(e, e, e), r);
}

// Quotient n/d without the IEEE divide's scaling and fix-up steps, for operands
// that are normal and finite (all sweep quantities are): FP64 multiplies by the
// refined reciprocal and applies one residual correction (within about one
// ulp); FP32 uses the hardware reciprocal directly (about two ulp, inside the
// FP32 sweep's error budget).
__device__ __forceinline__ double
fast_div(double n, double d)
{
  const double r = fast_rcp(d);
  const double q = n * r;
  return fma(r, fma(-d, q, n), q);
}

__device__ __forceinline__ float
fast_div(float n, float d)
{
  return n * __builtin_amdgcn_rcpf(d);
}

// Inverse TWDD weight 1/a for a = 1 - n/d clamped to [0.5, 1]: a < 0.5 <=> n > d/2
// and a > 1 <=> n < 0 (d > 0 whenever the upwind flux exceeds 1e-8).  Branch
// free: an infinite or NaN quotient (d - n == 0) is always replaced by a clamp.
template <typename R>
__device__ __forceinline__ R
inv_weight(R n, R d)
{
  const R dn = d - n;
  R r = fast_div(d, dn);
  r = n < R(0) ? R(1) : r;
  r = dn <= n ? R(2) : r;  // n >= d/2
  return r;
}

// The three inverse weights of one cell with a single reciprocal (FP64):
// 1/Da = Db Dc / (Da Db Dc) with Da = da - na etc. (batch inversion).  A weight
// that is clamped or takes the 0.5 default contributes D = 1, so no zero or
// non-finite factor can reach the shared reciprocal; an unclamped D lies in
// (d/2, d] and is well conditioned.  Within a few ulp of three separate divides.
__device__ __forceinline__ void
inv_weights3(double na, double da, bool va, double nb, double db, bool vb, double nc, double dc,
             bool vc, double & ia, double & ib, double & ic)
{
  // na < 0 is tested once, as the negation of na >= 0 (all operands are finite);
  // na < da/2 as da - na > na, on the difference the unclamped weight needs.
  const bool pa = na >= 0.0, pb = nb >= 0.0, pc = nc >= 0.0;
  const double ea = da - na, eb = db - nb, ec = dc - nc;
  const bool ua = va && pa && ea > na;
  const bool ub = vb && pb && eb > nb;
  const bool uc = vc && pc && ec > nc;
  const double Da = ua ? ea : 1.0;
  const double Db = ub ? eb : 1.0;
  const double Dc = uc ? ec : 1.0;
  const double Dbc = Db * Dc, P = Da * Dbc;
  const double r = fast_rcp(P);
  const double rDa = r * Da;
  const double ra = r * Dbc, rb = rDa * Dc, rc = rDa * Db;
  ia = ua ? da * ra : (va && !pa ? 1.0 : 2.0);
  ib = ub ? db * rb : (vb && !pb ? 1.0 : 2.0);
  ic = uc ? dc * rc : (vc && !pc ? 1.0 : 2.0);
}

// Straight-line TWDD3D: the three weights are always formed and the reference's
// "upwind flux > 1e-8" test selects between them and the 0.5 default, so a lane
// can interleave the solves of its directions.
template <typename R>
__device__ __forceinline__ R
twdd3d_fast(R qV, R sV, R mA, R eB, R xC, R & x, R & y, R & z)
{
  // n_a = th qV + th (eB y + xC z) + mA x = T + (1 - th) mA x with
  // T = th (qV + mA x + eB y + xC z), and likewise for b and c.
  const R th = R(0.9), om = R(0.1);
  const R mAx = mA * x, eBy = eB * y, xCz = xC * z;
  const R T = th * (qV + mAx + eBy + xCz);
  const R two_s = sV + R(2) * (mA + eB + xC);
  const R na = om * mAx + T, da = (two_s - R(2) * mA) * x;
  const R nb = om * eBy + T, db = (two_s - R(2) * eB) * y;
  const R nc = om * xCz + T, dc = (two_s - R(2) * xC) * z;
  R ia, ib, ic;
  if constexpr (sizeof(R) == 8)
  {
    inv_weights3(na, da, x > R(1e-8), nb, db, y > R(1e-8), nc, dc, z > R(1e-8), ia, ib, ic);
  }
  else
  {
    ia = inv_weight(na, da);
    ib = inv_weight(nb, db);
    ic = inv_weight(nc, dc);
    ia = x > R(1e-8) ? ia : R(2);
    ib = y > R(1e-8) ? ib : R(2);
    ic = z > R(1e-8) ? ic : R(2);
  }
  const R wa = mA * ia, wb = eB * ib, wc = xC * ic;
  const R psi = fast_div(qV + wa * x + wb * y + wc * z, sV + wa + wb + wc);
  // psi / a - ((1 - a) / a) x, with 1/a = ia, as x + ia (psi - x).
  x += ia * (psi - x);
  y += ib * (psi - y);
  z += ic * (psi - z);
  return psi;
}

template <typename R>
__device__ __forceinline__ R
dd3d_exact(R q, R sig, R lx, R ly, R lz, R amu, R aeta, R axi, R & x, R & y, R & z)
{
#pragma clang fp contract(off)
  R psi = q + (R(2.0) * x * amu / lx) + (R(2.0) * y * aeta / ly) + (R(2.0) * z * axi / lz);
  psi /= (R(2.0) * amu / lx) + (R(2.0) * aeta / ly) + (R(2.0) * axi / lz) + sig;
  x = R(2.0) * psi - x;
  y = R(2.0) * psi - y;
  z = R(2.0) * psi - z;
  return psi;
}

template <typename R>
__device__ __forceinline__ R
dd3d_fast(R q, R sig, R cx, R cy, R cz, R & x, R & y, R & z)
{
  const R psi = fast_div(q + cx * x + cy * y + cz * z, cx + cy + cz + sig);
  x = R(2) * psi - x;
  y = R(2) * psi - y;
  z = R(2) * psi - z;
  return psi;
}

template <typename R>
__device__ __forceinline__ R
twdd2d_exact(R q, R sig, R lx, R ly, R amu, R aeta, R axi, R & x, R & y)
{
#pragma clang fp contract(off)
  const R theta = R(0.9), theta_s = R(0.9);
  const R A = ly;
  const R B = lx;
  const R V = lx * ly;
  R a = R(0.5);
  if (x > R(1e-8))
  {
    a = q * V * theta_s + (aeta * B * y) * theta + amu * A * x;
    a /= (sig * V + R(2.0) * aeta * B) * x;
    a = R(-1.0) * (a - R(1.0));
  }
  R b = R(0.5);
  if (y > R(1e-8))
  {
    b = q * V * theta_s + (amu * A * x + axi) * theta + aeta * B * y;
    b /= (sig * V + R(2.0) * amu * A) * y;
    b = R(-1.0) * (b - R(1.0));
  }
  a = a < R(0.5) ? R(0.5) : a;
  a = a > R(1.0) ? R(1.0) : a;
  b = b < R(0.5) ? R(0.5) : b;
  b = b > R(1.0) ? R(1.0) : b;
  R psi = q * V + (amu * A / a * x) + (aeta * B / b * y);
  psi /= sig * V + (amu * A / a) + (aeta * B / b);
  x = (psi / a) - (((R(1.0) - a) / a) * x);
  y = (psi / b) - (((R(1.0) - b) / b) * y);
  return psi;
}

template <typename R>
__device__ __forceinline__ R
twdd2d_fast(R qV, R sV, R mA, R eB, R axi, R & x, R & y)
{
  const R th = R(0.9);
  const R qVs = qV * th;
  const R mAx = mA * x, eBy = eB * y;
  R ia = inv_weight(qVs + eBy * th + mAx, (sV + R(2) * eB) * x);
  R ib = inv_weight(qVs + (mAx + axi) * th + eBy, (sV + R(2) * mA) * y);
  ia = x > R(1e-8) ? ia : R(2);
  ib = y > R(1e-8) ? ib : R(2);
  const R wa = mA * ia, wb = eB * ib;
  const R psi = fast_div(qV + wa * x + wb * y, sV + wa + wb);
  x += ia * (psi - x);
  y += ib * (psi - y);
  return psi;
}

template <typename R>
__device__ __forceinline__ R
dd2d_exact(R q, R sig, R lx, R ly, R amu, R aeta, R & x, R & y)
{
#pragma clang fp contract(off)
  R psi = q + (R(2.0) * x * amu / lx) + (R(2.0) * y * aeta / ly);
  psi /= (R(2.0) * amu / lx) + (R(2.0) * aeta / ly) + sig;
  x = R(2.0) * psi - x;
  y = R(2.0) * psi - y;
  return psi;
}

template <typename R>
__device__ __forceinline__ R
dd1d(R q, R sig, R lx, R amu, R & x)
{
#pragma clang fp contract(off)
  R psi = q + (R(2.0) * x * amu / lx);
  psi /= (R(2.0) * amu / lx) + sig;
  x = R(2.0) * psi - x;
  return psi;
}

} // namespace doma
