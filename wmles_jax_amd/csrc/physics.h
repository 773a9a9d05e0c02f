// Point-wise physics of the explicit step, shared by every kernel.
//
// Each function follows the reference formula cited beside it and keeps its
// association order, so that a translation unit built with -ffp-contract=off
// reproduces the reference's float64 values bit for bit (ghost fills, geometry) and
// the contracted units agree to round-off.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define WMX_HD __host__ __device__ __forceinline__

namespace wmx {

struct GasC {            // reference wmles/gas.py:10-22
  double gamma, R, Pr, Pr_t, mu_ref, T_ref, S_suth, cp;
};

struct Prim {            // w = [rho, u, v, w, p] plus T
  double r, u, v, w, p, T;
};

WMX_HD double sq(double a) { return a * a; }
WMX_HD double dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
  return (a0 * b0 + a1 * b1) + a2 * b2;
}

// Sutherland, reference gas.py:24-26.  (T/T_ref)^1.5 = t sqrt(t).
WMX_HD double mu_suth(const GasC& g, double T) {
  double t = T / g.T_ref;
  return ((g.mu_ref * (t * sqrt(t))) * (g.T_ref + g.S_suth)) / (T + g.S_suth);
}

// reference gas.py:28-33
WMX_HD Prim cons_to_prim(const GasC& g, double q0, double q1, double q2, double q3, double q4) {
  Prim w;
  w.r = q0;
  w.u = q1 / q0;
  w.v = q2 / q0;
  w.w = q3 / q0;
  double k = (w.u * w.u + w.v * w.v) + w.w * w.w;
  w.p = (g.gamma - 1.0) * (q4 - (0.5 * q0) * k);
  w.T = w.p / (q0 * g.R);
  return w;
}

// reference gas.py:35-38
WMX_HD void prim_to_cons(const GasC& g, double r, double u, double v, double w, double p,
                         double* q) {
  double k = (u * u + v * v) + w * w;
  q[0] = r;
  q[1] = r * u;
  q[2] = r * v;
  q[3] = r * w;
  q[4] = p / (g.gamma - 1.0) + (0.5 * r) * k;
}

// ---------------------------------------------------------------- boundary ghost states
// SlipWall, reference bc.py:38-47
WMX_HD void bc_slip(const GasC& g, const double* src, double nx, double ny, double nz, double* out) {
  Prim w = cons_to_prim(g, src[0], src[1], src[2], src[3], src[4]);
  double un = dot3(w.u, w.v, w.w, nx, ny, nz);
  double t = 2.0 * un;
  prim_to_cons(g, w.r, w.u - t * nx, w.v - t * ny, w.w - t * nz, w.p, out);
}

// NoSlipWall, reference bc.py:50-64 (Tw < 0: adiabatic)
WMX_HD void bc_noslip(const GasC& g, const double* src, double Tw, double* out) {
  Prim w = cons_to_prim(g, src[0], src[1], src[2], src[3], src[4]);
  double Tg = w.T;
  if (Tw >= 0.0) Tg = fmax(2.0 * Tw - w.T, 0.1 * Tw);
  double rho = w.p / (g.R * Tg);
  prim_to_cons(g, rho, -w.u, -w.v, -w.w, w.p, out);
}

// CharacteristicFarfield, reference bc.py:80-115.  side: 0 (n_out = -n) or 1.
WMX_HD void bc_farfield(const GasC& g, const double* src, const double* pinf, double nx, double ny,
                        double nz, int side, double* out) {
  Prim w = cons_to_prim(g, src[0], src[1], src[2], src[3], src[4]);
  double ri = pinf[0], ui = pinf[1], vi = pinf[2], wi = pinf[3], pi = pinf[4];
  double ci = sqrt(g.gamma * pi / ri);
  double n0 = side == 1 ? nx : -nx, n1 = side == 1 ? ny : -ny, n2 = side == 1 ? nz : -nz;
  double un = dot3(w.u, w.v, w.w, n0, n1, n2);
  double un_inf = dot3(ui, vi, wi, n0, n1, n2);
  double c = sqrt(g.gamma * g.R * w.T);
  double drho = w.r - ri, dp = w.p - pi;
  double dun = un - un_inf;
  double ap = (un + c > 0.0) ? dp + ri * ci * dun : 0.0;
  double am = (un - c > 0.0) ? dp - ri * ci * dun : 0.0;
  double dp_bc = 0.5 * (ap + am);
  double dun_bc = (ap - am) / (2.0 * ri * ci);
  bool outgoing = un > 0.0;
  double ci2 = ci * ci;
  double entropy = outgoing ? drho - dp / ci2 : 0.0;
  double dt0 = (w.u - ui) - dun * n0, dt1 = (w.v - vi) - dun * n1, dt2 = (w.w - wi) - dun * n2;
  if (!outgoing) { dt0 = 0.0; dt1 = 0.0; dt2 = 0.0; }
  double rho_bc = (ri + dp_bc / ci2) + entropy;
  double u0 = (ui + dun_bc * n0) + dt0, u1 = (vi + dun_bc * n1) + dt1, u2 = (wi + dun_bc * n2) + dt2;
  prim_to_cons(g, rho_bc, u0, u1, u2, pi + dp_bc, out);
}

// --------------------------------------------------------------------- inviscid fluxes
// The bulk face arithmetic is templated on R: R = double is the plain form; R = float
// is the FP32-bulk form, in which p and e_int arrive gauge-shifted (p - p_ref,
// e_int - e_ref) and the caller adds the shift's contribution back in FP64.

// Division and square root of the templated bulk arithmetic.  R = double: the IEEE
// operations (the plain form's parity depends on them).  R = float (device code): the
// hardware reciprocal / square root (v_rcp_f32, v_sqrt_f32, about 1 ulp), which is
// well inside the FP32-bulk form's error budget.
template <class R>
WMX_HD R fdiv(R a, R b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(R) == 4) return a * __builtin_amdgcn_rcpf(b);
#endif
  return a / b;
}
template <class R>
WMX_HD R fsqrt(R a) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(R) == 4) return __builtin_amdgcn_sqrtf(a);
#endif
  return sqrt(a);
}

// Primitive record of one cell as the face kernels read it: rho, u, v, w, p, e_int.
template <class R>
struct CellWT {
  R r, u, v, w, p, e;   // e = p / ((gamma-1) rho)
};
using CellW = CellWT<double>;

template <class R>
WMX_HD R dot3r(R a0, R a1, R a2, R b0, R b1, R b2) {
  return (a0 * b0 + a1 * b1) + a2 * b2;
}

// Kennedy-Gruber split flux incl. area, reference schemes.py:15-23.  *un = u_bar . S.
template <class R>
WMX_HD void kg_flux(const CellWT<R>& a, const CellWT<R>& b, R Sx, R Sy, R Sz, R* F, R* un_out) {
  const R h = R(0.5);
  R rho = h * (a.r + b.r);
  R u = h * (a.u + b.u), v = h * (a.v + b.v), w = h * (a.w + b.w);
  R p = h * (a.p + b.p);
  R ka = (a.u * a.u + a.v * a.v) + a.w * a.w;
  R kb = (b.u * b.u + b.v * b.v) + b.w * b.w;
  R ea = h * (((a.e + h * ka) + b.e) + h * kb);
  R un = dot3r(u, v, w, Sx, Sy, Sz);
  R m = rho * un;
  F[0] = m;
  F[1] = m * u + p * Sx;
  F[2] = m * v + p * Sy;
  F[3] = m * w + p * Sz;
  F[4] = m * ea + p * un;
  *un_out = un;
}

// KEEP two-point flux, reference schemes.py:25-55.
template <class R>
WMX_HD void keep_flux(const CellWT<R>& a, const CellWT<R>& b, R Sx, R Sy, R Sz, R* F, R* un_out) {
  const R h = R(0.5);
  R rho = h * (a.r + b.r);
  R u = h * (a.u + b.u), v = h * (a.v + b.v), w = h * (a.w + b.w);
  R p = h * (a.p + b.p);
  R un = dot3r(u, v, w, Sx, Sy, Sz);
  R m = rho * un;
  R K = m * (h * dot3r(a.u, a.v, a.w, b.u, b.v, b.w));
  R I = (m * h) * (a.e + b.e);
  R unL = dot3r(a.u, a.v, a.w, Sx, Sy, Sz);
  R unR = dot3r(b.u, b.v, b.w, Sx, Sy, Sz);
  R P = h * (unR * a.p + unL * b.p);
  F[0] = m;
  F[1] = m * u + p * Sx;
  F[2] = m * v + p * Sy;
  F[3] = m * w + p * Sz;
  F[4] = (K + I) + P;
  *un_out = un;
}

// Central split form of order 2 or 4, reference schemes.py:57-70 (KEP), 72-98 (KEEP).
// *un4 is the same order-2/4 combination of u_bar . S (for the gauge term).
template <bool KEEP, class R>
WMX_HD void central_flux(int order, const CellWT<R>& L2, const CellWT<R>& L, const CellWT<R>& Rc,
                         const CellWT<R>& R2, R Sx, R Sy, R Sz, R* F, R* un4) {
  R F0[5], u0;
  if (KEEP) keep_flux(L, Rc, Sx, Sy, Sz, F0, &u0); else kg_flux(L, Rc, Sx, Sy, Sz, F0, &u0);
  if (order == 2) {
    for (int c = 0; c < 5; ++c) F[c] = F0[c];
    *un4 = u0;
    return;
  }
  R Fa[5], Fb[5], ua, ub;
  if (KEEP) {
    keep_flux(L, R2, Sx, Sy, Sz, Fa, &ua);
    keep_flux(L2, Rc, Sx, Sy, Sz, Fb, &ub);
  } else {
    kg_flux(L, R2, Sx, Sy, Sz, Fa, &ua);
    kg_flux(L2, Rc, Sx, Sy, Sz, Fb, &ub);
  }
  const R c23 = R(2.0 / 3.0), c112 = R(1.0 / 12.0), two = R(2.0);
  for (int c = 0; c < 5; ++c) F[c] = two * (c23 * F0[c] - c112 * (Fa[c] + Fb[c]));
  *un4 = two * (c23 * u0 - c112 * (ua + ub));
}

// Limiters, reference schemes.py:100-105.
template <class R>
WMX_HD R albada(R a, R b) {
  const R eps = R(1e-12);
  return fdiv<R>(a * (b * b + eps) + b * (a * a + eps), (a * a + b * b) + R(2.0) * eps);
}
template <class R>
WMX_HD R sgn(R a) { return a > R(0) ? R(1) : (a < R(0) ? R(-1) : R(0)); }
template <class R>
WMX_HD R minmod(R a, R b) {
  return (R(0.5) * (sgn(a) + sgn(b))) * fmin(fabs(a), fabs(b));
}

// HLLC with Davis speeds and Toro's star state, times |S|; reference schemes.py:108-137.
// wl, wr: [rho, u, v, w, p] with the full pressure.
template <class R>
WMX_HD void hllc_flux(R gamma, const R* wl, const R* wr, R Sx, R Sy, R Sz, R* F) {
  R Sm = fsqrt<R>((Sx * Sx + Sy * Sy) + Sz * Sz);
  R nx = fdiv<R>(Sx, Sm), ny = fdiv<R>(Sy, Sm), nz = fdiv<R>(Sz, Sm);
  R rl = wl[0], ul = wl[1], vl = wl[2], zl = wl[3], pl = wl[4];
  R rr = wr[0], ur = wr[1], vr = wr[2], zr = wr[3], pr = wr[4];
  R unl = dot3r(ul, vl, zl, nx, ny, nz), unr = dot3r(ur, vr, zr, nx, ny, nz);
  R cl = fsqrt<R>(fdiv<R>(gamma * pl, rl)), cr = fsqrt<R>(fdiv<R>(gamma * pr, rr));
  R sL = fmin(unl - cl, unr - cr), sR = fmax(unl + cl, unr + cr);
  R ss = fdiv<R>(((pr - pl) + (rl * unl) * (sL - unl)) - (rr * unr) * (sR - unr),
                 rl * (sL - unl) - rr * (sR - unr));
  const R gm1 = gamma - R(1.0), h = R(0.5);
  R El = fdiv<R>(pl, gm1) + (h * rl) * ((ul * ul + vl * vl) + zl * zl);
  R Er = fdiv<R>(pr, gm1) + (h * rr) * ((ur * ur + vr * vr) + zr * zr);
  R r, u, v, z, p, un, E, s;
  int branch;  // 0: FL, 1: F*L, 2: F*R, 3: FR
  if (sL >= R(0)) branch = 0;
  else if (ss >= R(0)) branch = 1;
  else if (sR > R(0)) branch = 2;
  else branch = 3;
  if (branch <= 1) { r = rl; u = ul; v = vl; z = zl; p = pl; un = unl; E = El; s = sL; }
  else { r = rr; u = ur; v = vr; z = zr; p = pr; un = unr; E = Er; s = sR; }
  R m = r * un;
  R Fp[5] = {m, m * u + p * nx, m * v + p * ny, m * z + p * nz, (E + p) * un};
  if (branch == 1 || branch == 2) {
    R f = fdiv<R>(r * (s - un), s - ss);
    R e = fdiv<R>(E, r) + (ss - un) * (ss + fdiv<R>(p, r * (s - un)));
    R d = ss - un;
    R st[5] = {f, f * (u + d * nx), f * (v + d * ny), f * (z + d * nz), f * e};
    R U[5] = {r, r * u, r * v, r * z, E};
    for (int c = 0; c < 5; ++c) Fp[c] = Fp[c] + s * (st[c] - U[c]);
  }
  for (int c = 0; c < 5; ++c) F[c] = Fp[c] * Sm;
}

// MUSCL (primitive, index space) + positivity fallback + HLLC, reference schemes.py:140-152.
// p_ref is added back to the (gauge-shifted) reconstructed pressures before HLLC.
template <class R>
WMX_HD void muscl_hllc_flux(R gamma, int limiter, const CellWT<R>& L2, const CellWT<R>& L,
                            const CellWT<R>& Rc, const CellWT<R>& R2, R Sx, R Sy, R Sz, R p_ref,
                            R* F) {
  const R l2[5] = {L2.r, L2.u, L2.v, L2.w, L2.p};
  const R l1[5] = {L.r, L.u, L.v, L.w, L.p};
  const R r1[5] = {Rc.r, Rc.u, Rc.v, Rc.w, Rc.p};
  const R r2[5] = {R2.r, R2.u, R2.v, R2.w, R2.p};
  R wl[5], wr[5];
  const R h = R(0.5);
  for (int c = 0; c < 5; ++c) {
    R a = l1[c] - l2[c], b = r1[c] - l1[c], d = r2[c] - r1[c];
    R la = limiter == 0 ? albada(a, b) : minmod(a, b);
    R lb = limiter == 0 ? albada(b, d) : minmod(b, d);
    wl[c] = l1[c] + h * la;
    wr[c] = r1[c] - h * lb;
  }
  wl[4] = wl[4] + p_ref;
  wr[4] = wr[4] + p_ref;
  if (wl[0] <= R(0) || wl[4] <= R(0) || wr[0] <= R(0) || wr[4] <= R(0)) {
    for (int c = 0; c < 5; ++c) { wl[c] = l1[c]; wr[c] = r1[c]; }
    wl[4] = wl[4] + p_ref;
    wr[4] = wr[4] + p_ref;
  }
  hllc_flux(gamma, wl, wr, Sx, Sy, Sz, F);
}

// ---------------------------------------------------------------------- viscous flux
// Cell record the viscous flux reads: grad [u,v,w,T] (row-major g[v][c]), mu_eff, k_eff.
template <class R>
struct CellGT {
  R g[12];
  R mu, k;
};
using CellG = CellGT<double>;

// reference viscous.py:28-51 for one face; phi = [u, v, w, T]; adds into F.
// e (unit cell-to-cell vector) and idist = 1 / dist come from FP64 centroids (the four
// differences phi_R - phi_L are scaled by one reciprocal: within an ulp of the reference's
// division).
template <class R>
WMX_HD void viscous_face_e(const CellGT<R>& a, const CellGT<R>& b, const R* phiL, const R* phiR,
                           R e0, R e1, R e2, R idist, R Sx, R Sy, R Sz, R* F) {
  const R h = R(0.5);
  R gf[12];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    R g0 = h * (a.g[3 * v + 0] + b.g[3 * v + 0]);
    R g1 = h * (a.g[3 * v + 1] + b.g[3 * v + 1]);
    R g2 = h * (a.g[3 * v + 2] + b.g[3 * v + 2]);
    R corr = (phiR[v] - phiL[v]) * idist - ((g0 * e0 + g1 * e1) + g2 * e2);
    gf[3 * v + 0] = g0 + corr * e0;
    gf[3 * v + 1] = g1 + corr * e1;
    gf[3 * v + 2] = g2 + corr * e2;
  }
  R muf = h * (a.mu + b.mu), kf = h * (a.k + b.k);
  R div = (gf[0] + gf[4]) + gf[8];
  R tdiv = R(2.0 / 3.0) * div;
  R t00 = muf * ((gf[0] + gf[0]) - tdiv), t11 = muf * ((gf[4] + gf[4]) - tdiv),
    t22 = muf * ((gf[8] + gf[8]) - tdiv);
  R t01 = muf * (gf[1] + gf[3]), t02 = muf * (gf[2] + gf[6]), t12 = muf * (gf[5] + gf[7]);
  R ts0 = (t00 * Sx + t01 * Sy) + t02 * Sz;
  R ts1 = (t01 * Sx + t11 * Sy) + t12 * Sz;
  R ts2 = (t02 * Sx + t12 * Sy) + t22 * Sz;
  R dTdn = (gf[9] * Sx + gf[10] * Sy) + gf[11] * Sz;
  R uf0 = h * (phiL[0] + phiR[0]), uf1 = h * (phiL[1] + phiR[1]), uf2 = h * (phiL[2] + phiR[2]);
  F[0] = F[0] + R(0.0);
  F[1] = F[1] + (-ts0);
  F[2] = F[2] + (-ts1);
  F[3] = F[3] + (-ts2);
  F[4] = F[4] + ((-((uf0 * ts0 + uf1 * ts1) + uf2 * ts2)) - kf * dTdn);
}

WMX_HD void viscous_face(const CellG& a, const CellG& b, const double* phiL, const double* phiR,
                         const double* xcL, const double* xcR, double Sx, double Sy, double Sz,
                         double* F) {
  double dx0 = xcR[0] - xcL[0], dx1 = xcR[1] - xcL[1], dx2 = xcR[2] - xcL[2];
  double idist = 1.0 / sqrt((dx0 * dx0 + dx1 * dx1) + dx2 * dx2);
  viscous_face_e<double>(a, b, phiL, phiR, dx0 * idist, dx1 * idist, dx2 * idist, idist, Sx, Sy, Sz, F);
}

// Plain-form KEP/KEEP wrappers used by the FP64 kernels (no gauge).
template <bool KEEP>
WMX_HD void central_flux(int order, const CellW& L2, const CellW& L, const CellW& R, const CellW& R2,
                         double Sx, double Sy, double Sz, double* F) {
  double un4;
  central_flux<KEEP, double>(order, L2, L, R, R2, Sx, Sy, Sz, F, &un4);
}
WMX_HD void muscl_hllc_flux(double gamma, int limiter, const CellW& L2, const CellW& L,
                            const CellW& R, const CellW& R2, double Sx, double Sy, double Sz,
                            double* F) {
  muscl_hllc_flux<double>(gamma, limiter, L2, L, R, R2, Sx, Sy, Sz, 0.0, F);
}

// ------------------------------------------------------------------------------ SGS
// g: velocity gradient g[i][j] = du_i/dx_j row-major (first 9 of CellG.g).
// reference sgs.py:14-18
template <class R>
WMX_HD R nut_smagorinsky(const R* g, R delta, R Cs) {
  R ss = R(0.0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R s = R(0.5) * (g[3 * i + j] + g[3 * j + i]);
      ss = ss + s * s;
    }
  R cd = Cs * delta;
  return (cd * cd) * fsqrt<R>(R(2.0) * ss);
}

// S:S and Sd:Sd of the WALE model for the gradient g.
template <class R>
WMX_HD void wale_invariants(const R* g, R& ss, R& sdsd) {
  ss = R(0.0);
  R g2[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R s = R(0.5) * (g[3 * i + j] + g[3 * j + i]);
      ss = ss + s * s;
      g2[3 * i + j] = (g[3 * i + 0] * g[0 + j] + g[3 * i + 1] * g[3 + j]) + g[3 * i + 2] * g[6 + j];
    }
  R tr = (g2[0] + g2[4]) + g2[8];
  sdsd = R(0.0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      R sd = R(0.5) * (g2[3 * i + j] + g2[3 * j + i]) - (i == j ? fdiv<R>(tr, R(3.0)) : R(0.0));
      sdsd = sdsd + sd * sd;
    }
}

// reference sgs.py:21-30.  The FP64 form evaluates the reference expression as written.
// In float the numerator (Sd:Sd)^(3/2) grows like |g|^6 and overflows beyond
// |g| ~ 2.6e6 1/s (near-wall shear of millimetre channels reaches 1e7), so the float
// form evaluates the model on g/|g| (|g| the Frobenius norm) and scales back by |g|: the
// model is homogeneous of degree one in g, the regulariser 1e-30 becomes 1e-30/|g|^5.
template <class R>
WMX_HD R nut_wale(const R* g, R delta, R Cw) {
  R cd = Cw * delta;
  if constexpr (sizeof(R) == 4) {
    R n2 = R(0.0);
#pragma unroll
    for (int k = 0; k < 9; ++k) n2 = n2 + g[k] * g[k];
    if (!(n2 > R(0.0))) return R(0.0);
    R gn = fsqrt<R>(n2);
    R rg = fdiv<R>(R(1.0), gn);
    R gs[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) gs[k] = g[k] * rg;
    R ss, sdsd;
    wale_invariants<R>(gs, ss, sdsd);
    R rg2 = rg * rg;
    R num = (cd * cd) * gn * (sdsd * fsqrt<R>(sdsd));
    R den = ((ss * ss) * fsqrt<R>(ss) + fsqrt<R>(fsqrt<R>(sdsd)) * sdsd) + R(1e-30) * ((rg2 * rg2) * rg);
    return fdiv<R>(num, den);
  } else {
    R ss, sdsd;
    wale_invariants<R>(g, ss, sdsd);
    R num = (cd * cd) * (sdsd * fsqrt<R>(sdsd));
    R den = ((ss * ss) * fsqrt<R>(ss) + fsqrt<R>(fsqrt<R>(sdsd)) * sdsd) + R(1e-30);
    return fdiv<R>(num, den);
  }
}

// reference sgs.py:33-42
template <class R>
WMX_HD R nut_vreman(const R* g, R delta, R Cs) {
  R c = R(2.5) * (Cs * Cs);
  R d2 = delta * delta;
  R b[9];
  R aa = R(0.0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      b[3 * i + j] = d2 * ((g[3 * i + 0] * g[3 * j + 0] + g[3 * i + 1] * g[3 * j + 1]) +
                           g[3 * i + 2] * g[3 * j + 2]);
      aa = aa + g[3 * i + j] * g[3 * i + j];
    }
  R B = ((((b[0] * b[4] - b[1] * b[1]) + b[0] * b[8]) - b[2] * b[2]) + b[4] * b[8]) - b[5] * b[5];
  return c * fsqrt<R>(fdiv<R>(fmax(B, R(0.0)), aa + R(1e-30)));
}

// reference viscous.py:22-25
template <class R>
WMX_HD R ducros(const R* g) {
  R div = (g[0] + g[4]) + g[8];
  R a = g[7] - g[5], b = g[2] - g[6], c = g[3] - g[1];
  R w2 = (a * a + b * b) + c * c;
  R d2 = div * div;
  return fdiv<R>(d2, (d2 + w2) + R(1e-30));
}

// ------------------------------------------------------------------------ wall model
struct WallModelC {      // reference wall_model.py:15-17 + host-evaluated constants
  double kappa, B, ekb, pr13, pr3, beta;
  int n_iter;
};

// Spalding u_tau by safeguarded fixed-count Newton, reference wall_model.py:19-40.  The
// iteration map depends on up alone, so once an update returns its input exactly the
// remaining iterations would repeat it: stopping there gives the fixed-count result bit
// for bit.
WMX_HD double spalding_utau(const WallModelC& m, double h, double U, double rho_w, double mu_w) {
  double k = m.kappa, ekb = m.ekb;
  double Reh = ((h * U) * rho_w) / mu_w;
  double up = fmin(fmax(sqrt(Reh), 1.0), 25.0);
  for (int it = 0; it < m.n_iter; ++it) {
    double ku = k * up;
    double ex = exp(ku);
    double poly2 = ((ex - 1.0) - ku) - 0.5 * (ku * ku);
    double E = ekb * (poly2 - (ku * ku * ku) / 6.0);
    double F = (up + E) - Reh / up;
    double dF = (1.0 + (ekb * k) * poly2) + Reh / (up * up);
    double nw = up - F / dF;
    const double upn = fmin(fmax(nw, 0.5 * up), 100.0);
    if (upn == up) break;
    up = upn;
  }
  return U / up;
}

// reference wall_model.py:42-61.  Tw < 0: adiabatic.  Returns tau_w, q_w.
WMX_HD void equilibrium_wall(const GasC& g, const WallModelC& m, double h, double U, double T_h,
                             double p_h, double Tw_in, double* tau_w, double* q_w) {
  double Pr = g.Pr;
  U = fmax(U, 1e-8);
  bool adiabatic = Tw_in < 0.0;
  double T_w = adiabatic ? T_h + (m.pr13 * (U * U)) / (2.0 * g.cp) : Tw_in + 0.0 * T_h;
  double rho_w = p_h / (g.R * T_w);
  double mu_w = mu_suth(g, T_w);
  double ut = spalding_utau(m, h, U, rho_w, mu_w);
  double tau = rho_w * (ut * ut);
  *tau_w = tau;
  if (adiabatic) { *q_w = 0.0 * tau; return; }
  double yp = fmax(((h * ut) * rho_w) / mu_w, 1e-6);
  double py = Pr * yp;
  double py2 = py * py;
  double Gam = fmax((0.01 * (py2 * py2)) / (1.0 + (5.0 * m.pr3) * yp), 1e-20);
  double Tp = (Pr * yp) * exp(-Gam) + (2.12 * log(1.0 + yp) + m.beta) * exp(-1.0 / Gam);
  *q_w = (((rho_w * g.cp) * ut) * (T_w - T_h)) / Tp;
}

}  // namespace wmx
