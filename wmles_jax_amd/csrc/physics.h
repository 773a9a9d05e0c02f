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
// Primitive record of one cell as the face kernels read it: rho, u, v, w, p, e_int.
struct CellW {
  double r, u, v, w, p, e;   // e = p / ((gamma-1) rho)
};

// Kennedy-Gruber split flux incl. area, reference schemes.py:15-23.
WMX_HD void kg_flux(const CellW& a, const CellW& b, double Sx, double Sy, double Sz, double* F) {
  double rho = 0.5 * (a.r + b.r);
  double u = 0.5 * (a.u + b.u), v = 0.5 * (a.v + b.v), w = 0.5 * (a.w + b.w);
  double p = 0.5 * (a.p + b.p);
  double ka = (a.u * a.u + a.v * a.v) + a.w * a.w;
  double kb = (b.u * b.u + b.v * b.v) + b.w * b.w;
  double ea = 0.5 * (((a.e + 0.5 * ka) + b.e) + 0.5 * kb);
  double un = dot3(u, v, w, Sx, Sy, Sz);
  double m = rho * un;
  F[0] = m;
  F[1] = m * u + p * Sx;
  F[2] = m * v + p * Sy;
  F[3] = m * w + p * Sz;
  F[4] = m * ea + p * un;
}

// KEEP two-point flux, reference schemes.py:25-55.
WMX_HD void keep_flux(const CellW& a, const CellW& b, double Sx, double Sy, double Sz, double* F) {
  double rho = 0.5 * (a.r + b.r);
  double u = 0.5 * (a.u + b.u), v = 0.5 * (a.v + b.v), w = 0.5 * (a.w + b.w);
  double p = 0.5 * (a.p + b.p);
  double un = dot3(u, v, w, Sx, Sy, Sz);
  double m = rho * un;
  double K = m * (0.5 * dot3(a.u, a.v, a.w, b.u, b.v, b.w));
  double I = (m * 0.5) * (a.e + b.e);
  double unL = dot3(a.u, a.v, a.w, Sx, Sy, Sz);
  double unR = dot3(b.u, b.v, b.w, Sx, Sy, Sz);
  double P = 0.5 * (unR * a.p + unL * b.p);
  F[0] = m;
  F[1] = m * u + p * Sx;
  F[2] = m * v + p * Sy;
  F[3] = m * w + p * Sz;
  F[4] = (K + I) + P;
}

// Central split form of order 2 or 4, reference schemes.py:57-70 (KEP), 72-98 (KEEP).
template <bool KEEP>
WMX_HD void central_flux(int order, const CellW& L2, const CellW& L, const CellW& R,
                         const CellW& R2, double Sx, double Sy, double Sz, double* F) {
  double F0[5];
  if (KEEP) keep_flux(L, R, Sx, Sy, Sz, F0); else kg_flux(L, R, Sx, Sy, Sz, F0);
  if (order == 2) {
    for (int c = 0; c < 5; ++c) F[c] = F0[c];
    return;
  }
  double Fa[5], Fb[5];
  if (KEEP) {
    keep_flux(L, R2, Sx, Sy, Sz, Fa);
    keep_flux(L2, R, Sx, Sy, Sz, Fb);
  } else {
    kg_flux(L, R2, Sx, Sy, Sz, Fa);
    kg_flux(L2, R, Sx, Sy, Sz, Fb);
  }
  for (int c = 0; c < 5; ++c)
    F[c] = 2.0 * ((2.0 / 3.0) * F0[c] - (1.0 / 12.0) * (Fa[c] + Fb[c]));
}

// Limiters, reference schemes.py:100-105.
WMX_HD double albada(double a, double b) {
  const double eps = 1e-12;
  return (a * (b * b + eps) + b * (a * a + eps)) / ((a * a + b * b) + 2.0 * eps);
}
WMX_HD double sgn(double a) { return a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0); }
WMX_HD double minmod(double a, double b) {
  return (0.5 * (sgn(a) + sgn(b))) * fmin(fabs(a), fabs(b));
}

// HLLC with Davis speeds and Toro's star state, times |S|; reference schemes.py:108-137.
// wl, wr: [rho, u, v, w, p].
WMX_HD void hllc_flux(double gamma, const double* wl, const double* wr, double Sx, double Sy,
                      double Sz, double* F) {
  double Sm = sqrt((Sx * Sx + Sy * Sy) + Sz * Sz);
  double nx = Sx / Sm, ny = Sy / Sm, nz = Sz / Sm;
  double rl = wl[0], ul = wl[1], vl = wl[2], zl = wl[3], pl = wl[4];
  double rr = wr[0], ur = wr[1], vr = wr[2], zr = wr[3], pr = wr[4];
  double unl = dot3(ul, vl, zl, nx, ny, nz), unr = dot3(ur, vr, zr, nx, ny, nz);
  double cl = sqrt(gamma * pl / rl), cr = sqrt(gamma * pr / rr);
  double sL = fmin(unl - cl, unr - cr), sR = fmax(unl + cl, unr + cr);
  double ss = (((pr - pl) + (rl * unl) * (sL - unl)) - (rr * unr) * (sR - unr)) /
              (rl * (sL - unl) - rr * (sR - unr));
  double El = pl / (gamma - 1.0) + (0.5 * rl) * ((ul * ul + vl * vl) + zl * zl);
  double Er = pr / (gamma - 1.0) + (0.5 * rr) * ((ur * ur + vr * vr) + zr * zr);
  double r, u, v, z, p, un, E, s;
  int branch;  // 0: FL, 1: F*L, 2: F*R, 3: FR
  if (sL >= 0.0) branch = 0;
  else if (ss >= 0.0) branch = 1;
  else if (sR > 0.0) branch = 2;
  else branch = 3;
  if (branch <= 1) { r = rl; u = ul; v = vl; z = zl; p = pl; un = unl; E = El; s = sL; }
  else { r = rr; u = ur; v = vr; z = zr; p = pr; un = unr; E = Er; s = sR; }
  double m = r * un;
  double Fp[5] = {m, m * u + p * nx, m * v + p * ny, m * z + p * nz, (E + p) * un};
  if (branch == 1 || branch == 2) {
    double f = r * (s - un) / (s - ss);
    double e = E / r + (ss - un) * (ss + p / (r * (s - un)));
    double d = ss - un;
    double st[5] = {f, f * (u + d * nx), f * (v + d * ny), f * (z + d * nz), f * e};
    double U[5] = {r, r * u, r * v, r * z, E};
    for (int c = 0; c < 5; ++c) Fp[c] = Fp[c] + s * (st[c] - U[c]);
  }
  for (int c = 0; c < 5; ++c) F[c] = Fp[c] * Sm;
}

// MUSCL (primitive, index space) + positivity fallback + HLLC, reference schemes.py:140-152.
WMX_HD void muscl_hllc_flux(double gamma, int limiter, const CellW& L2, const CellW& L,
                            const CellW& R, const CellW& R2, double Sx, double Sy, double Sz,
                            double* F) {
  const double l2[5] = {L2.r, L2.u, L2.v, L2.w, L2.p};
  const double l1[5] = {L.r, L.u, L.v, L.w, L.p};
  const double r1[5] = {R.r, R.u, R.v, R.w, R.p};
  const double r2[5] = {R2.r, R2.u, R2.v, R2.w, R2.p};
  double wl[5], wr[5];
  for (int c = 0; c < 5; ++c) {
    double a = l1[c] - l2[c], b = r1[c] - l1[c], d = r2[c] - r1[c];
    double la = limiter == 0 ? albada(a, b) : minmod(a, b);
    double lb = limiter == 0 ? albada(b, d) : minmod(b, d);
    wl[c] = l1[c] + 0.5 * la;
    wr[c] = r1[c] - 0.5 * lb;
  }
  if (wl[0] <= 0.0 || wl[4] <= 0.0 || wr[0] <= 0.0 || wr[4] <= 0.0) {
    for (int c = 0; c < 5; ++c) { wl[c] = l1[c]; wr[c] = r1[c]; }
  }
  hllc_flux(gamma, wl, wr, Sx, Sy, Sz, F);
}

// ---------------------------------------------------------------------- viscous flux
// Cell record the viscous flux reads: grad [u,v,w,T] (row-major g[v][c]), mu_eff, k_eff.
struct CellG {
  double g[12];
  double mu, k;
};

// reference viscous.py:28-51 for one face; phi = [u, v, w, T]; adds into F.
WMX_HD void viscous_face(const CellG& a, const CellG& b, const double* phiL, const double* phiR,
                         const double* xcL, const double* xcR, double Sx, double Sy, double Sz,
                         double* F) {
  double dx0 = xcR[0] - xcL[0], dx1 = xcR[1] - xcL[1], dx2 = xcR[2] - xcL[2];
  double dist = sqrt((dx0 * dx0 + dx1 * dx1) + dx2 * dx2);
  double e0 = dx0 / dist, e1 = dx1 / dist, e2 = dx2 / dist;
  double gf[12];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    double g0 = 0.5 * (a.g[3 * v + 0] + b.g[3 * v + 0]);
    double g1 = 0.5 * (a.g[3 * v + 1] + b.g[3 * v + 1]);
    double g2 = 0.5 * (a.g[3 * v + 2] + b.g[3 * v + 2]);
    double corr = (phiR[v] - phiL[v]) / dist - ((g0 * e0 + g1 * e1) + g2 * e2);
    gf[3 * v + 0] = g0 + corr * e0;
    gf[3 * v + 1] = g1 + corr * e1;
    gf[3 * v + 2] = g2 + corr * e2;
  }
  double muf = 0.5 * (a.mu + b.mu), kf = 0.5 * (a.k + b.k);
  double div = (gf[0] + gf[4]) + gf[8];
  double tdiv = (2.0 / 3.0) * div;
  double t00 = muf * ((gf[0] + gf[0]) - tdiv), t11 = muf * ((gf[4] + gf[4]) - tdiv),
         t22 = muf * ((gf[8] + gf[8]) - tdiv);
  double t01 = muf * (gf[1] + gf[3]), t02 = muf * (gf[2] + gf[6]), t12 = muf * (gf[5] + gf[7]);
  double ts0 = (t00 * Sx + t01 * Sy) + t02 * Sz;
  double ts1 = (t01 * Sx + t11 * Sy) + t12 * Sz;
  double ts2 = (t02 * Sx + t12 * Sy) + t22 * Sz;
  double dTdn = (gf[9] * Sx + gf[10] * Sy) + gf[11] * Sz;
  double uf0 = 0.5 * (phiL[0] + phiR[0]), uf1 = 0.5 * (phiL[1] + phiR[1]),
         uf2 = 0.5 * (phiL[2] + phiR[2]);
  F[0] = F[0] + 0.0;
  F[1] = F[1] + (-ts0);
  F[2] = F[2] + (-ts1);
  F[3] = F[3] + (-ts2);
  F[4] = F[4] + ((-((uf0 * ts0 + uf1 * ts1) + uf2 * ts2)) - kf * dTdn);
}

// ------------------------------------------------------------------------------ SGS
// g: velocity gradient g[i][j] = du_i/dx_j row-major (first 9 of CellG.g).
// reference sgs.py:14-18
WMX_HD double nut_smagorinsky(const double* g, double delta, double Cs) {
  double ss = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.5 * (g[3 * i + j] + g[3 * j + i]);
      ss = ss + s * s;
    }
  double cd = Cs * delta;
  return (cd * cd) * sqrt(2.0 * ss);
}

// reference sgs.py:21-30
WMX_HD double nut_wale(const double* g, double delta, double Cw) {
  double ss = 0.0;
  double g2[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.5 * (g[3 * i + j] + g[3 * j + i]);
      ss = ss + s * s;
      g2[3 * i + j] = (g[3 * i + 0] * g[0 + j] + g[3 * i + 1] * g[3 + j]) + g[3 * i + 2] * g[6 + j];
    }
  double tr = (g2[0] + g2[4]) + g2[8];
  double sdsd = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double sd = 0.5 * (g2[3 * i + j] + g2[3 * j + i]) - (i == j ? tr / 3.0 : 0.0);
      sdsd = sdsd + sd * sd;
    }
  double cd = Cw * delta;
  double num = (cd * cd) * (sdsd * sqrt(sdsd));
  double den = ((ss * ss) * sqrt(ss) + sqrt(sqrt(sdsd)) * sdsd) + 1e-30;
  return num / den;
}

// reference sgs.py:33-42
WMX_HD double nut_vreman(const double* g, double delta, double Cs) {
  double c = 2.5 * (Cs * Cs);
  double d2 = delta * delta;
  double b[9];
  double aa = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      b[3 * i + j] = d2 * ((g[3 * i + 0] * g[3 * j + 0] + g[3 * i + 1] * g[3 * j + 1]) +
                           g[3 * i + 2] * g[3 * j + 2]);
      aa = aa + g[3 * i + j] * g[3 * i + j];
    }
  double B = ((((b[0] * b[4] - b[1] * b[1]) + b[0] * b[8]) - b[2] * b[2]) + b[4] * b[8]) - b[5] * b[5];
  return c * sqrt(fmax(B, 0.0) / (aa + 1e-30));
}

// reference viscous.py:22-25
WMX_HD double ducros(const double* g) {
  double div = (g[0] + g[4]) + g[8];
  double w2 = (sq(g[7] - g[5]) + sq(g[2] - g[6])) + sq(g[3] - g[1]);
  double d2 = div * div;
  return d2 / ((d2 + w2) + 1e-30);
}

// ------------------------------------------------------------------------ wall model
struct WallModelC {      // reference wall_model.py:15-17 + host-evaluated constants
  double kappa, B, ekb, pr13, pr3, beta;
  int n_iter;
};

// Spalding u_tau by safeguarded fixed-count Newton, reference wall_model.py:19-40.
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
    up = fmin(fmax(nw, 0.5 * up), 100.0);
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
