#include "host_math.hpp"

#include <cmath>

namespace rcwa_host {

WaveVectors wave_vectors(double Lx, double Ly, const Incidence& inc, int hx, int hy,
                         const Media& m) {
  WaveVectors w;
  const cplx n_inc = std::sqrt(m.er_ref * m.ur_ref);  // plain sqrt (rcwa.cpp:91)
  const double st = std::sin(inc.theta), ct = std::cos(inc.theta);
  const double sp = std::sin(inc.phi), cp = std::cos(inc.phi);
  w.k_inc[0] = n_inc * st * cp;                       // rcwa.cpp:99-103
  w.k_inc[1] = n_inc * st * sp;
  w.k_inc[2] = n_inc * ct;
  const int P = 2 * hx + 1, Q = 2 * hy + 1, PQ = P * Q;
  w.Kx.resize(PQ);
  w.Ky.resize(PQ);
  w.Kz_ref.resize(PQ);
  w.Kz_trn.resize(PQ);
  const cplx er_r = std::conj(m.ur_ref) * std::conj(m.er_ref);
  const cplx er_t = std::conj(m.ur_trn) * std::conj(m.er_trn);
  for (int ix = 0; ix < P; ++ix) {
    const cplx kx = w.k_inc[0] - (double)(ix - hx) * (inc.lambda0 / Lx);  // rcwa.cpp:106-112
    for (int iy = 0; iy < Q; ++iy) {
      const cplx ky = w.k_inc[1] - (double)(iy - hy) * (inc.lambda0 / Ly);
      const int idx = ix * Q + iy;                     // MeshGrid, column-major flatten
      w.Kx[idx] = kx;
      w.Ky[idx] = ky;
      const cplx k2 = kx * kx + ky * ky;
      w.Kz_ref[idx] = -std::conj(usqrt(er_r - k2));   // rcwa.cpp:134
      w.Kz_trn[idx] = std::conj(usqrt(er_t - k2));    // rcwa.cpp:135
    }
  }
  return w;
}

std::vector<cplx> gap_v0_diagonals(const std::vector<cplx>& Kx, const std::vector<cplx>& Ky) {
  const int N = (int)Kx.size();
  std::vector<cplx> d(4 * N);
  for (int i = 0; i < N; ++i) {
    const cplx kz = std::conj(usqrt(1.0 - Kx[i] * Kx[i] - Ky[i] * Ky[i]));  // rcwa.cpp:179
    const cplx lam = cplx(0, 1) * kz;                                       // rcwa.cpp:214
    d[0 * N + i] = (Kx[i] * Ky[i]) / lam;
    d[1 * N + i] = (1.0 - Kx[i] * Kx[i]) / lam;
    d[2 * N + i] = (Ky[i] * Ky[i] - 1.0) / lam;
    d[3 * N + i] = (-(Kx[i] * Ky[i])) / lam;
  }
  return d;
}

namespace {
struct M2 {
  cplx a, b, c, d;  // [[a b][c d]]
};
inline M2 mul(const M2& x, const M2& y) {
  return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c,
          x.c * y.b + x.d * y.d};
}
inline M2 inv(const M2& x) {
  const cplx det = x.a * x.d - x.b * x.c;
  return {x.d / det, -x.b / det, -x.c / det, x.a / det};
}
inline M2 add(const M2& x, const M2& y, double s = 1.0) {
  return {x.a + s * y.a, x.b + s * y.b, x.c + s * y.c, x.d + s * y.d};
}
inline M2 scal(const M2& x, cplx s) { return {x.a * s, x.b * s, x.c * s, x.d * s}; }
}  // namespace

std::vector<cplx> end_smatrix_diagonals(int which, const std::vector<cplx>& Kx,
                                        const std::vector<cplx>& Ky, const std::vector<cplx>& Kz,
                                        cplx er, cplx ur) {
  const int N = (int)Kx.size();
  std::vector<cplx> out(16 * (size_t)N);
  const cplx ue = ur * er;  // no conjugate (rcwa.cpp:362-363, 422-423)
  for (int m = 0; m < N; ++m) {
    const cplx kx = Kx[m], ky = Ky[m];
    const cplx kz0 = std::conj(usqrt(1.0 - kx * kx - ky * ky));
    const cplx lam0 = cplx(0, 1) * kz0;
    const M2 q0{kx * ky, 1.0 - kx * kx, ky * ky - 1.0, -(kx * ky)};
    const M2 qe = scal(M2{kx * ky, ue - kx * kx, ky * ky - ue, -(kx * ky)}, 1.0 / ur);
    const cplx lame = (which == 0 ? cplx(0, -1) : cplx(0, 1)) * Kz[m];  // rcwa.cpp:377, 437
    // V0^-1 V_end restricted to harmonic m: (q0/lam0)^-1 (qe/lame)
    const M2 vv = scal(mul(inv(q0), qe), lam0 / lame);
    const M2 I{1.0, 0.0, 0.0, 1.0};
    const M2 A = add(I, vv), B = add(I, vv, -1.0);
    const M2 Ai = inv(A);
    const M2 AiB = mul(Ai, B);
    const M2 sa = scal(AiB, -1.0);                        // -A^-1 B
    const M2 sb = scal(Ai, 2.0);                          // 2 A^-1
    const M2 sc = scal(add(A, mul(B, AiB), -1.0), 0.5);   // (A - B A^-1 B)/2
    const M2 sd = mul(B, Ai);                             // B A^-1
    const M2 blk[4] = {which == 0 ? sa : sd, which == 0 ? sb : sc, which == 0 ? sc : sb,
                       which == 0 ? sd : sa};
    for (int k = 0; k < 4; ++k) {
      out[(size_t)(4 * k + 0) * N + m] = blk[k].a;
      out[(size_t)(4 * k + 1) * N + m] = blk[k].b;
      out[(size_t)(4 * k + 2) * N + m] = blk[k].c;
      out[(size_t)(4 * k + 3) * N + m] = blk[k].d;
    }
  }
  return out;
}

Efficiencies efficiencies(double theta, double pte, double ptm, const WaveVectors& wv,
                          const Media& m, const cplx* s11c, const cplx* s21c) {
  const int PQ = (int)wv.Kx.size();
  const cplx* k = wv.k_inc;
  cplx aTE[3];
  if (std::fabs(theta) < 1e-6) {                     // rcwa.cpp:483-487
    aTE[0] = 0.0; aTE[1] = 1.0; aTE[2] = 0.0;
  } else {
    aTE[0] = -k[1]; aTE[1] = k[0]; aTE[2] = 0.0;
  }
  auto normalize = [](cplx* v) {
    const double nn = std::sqrt(std::norm(v[0]) + std::norm(v[1]) + std::norm(v[2]));
    for (int i = 0; i < 3; ++i) v[i] /= nn;
  };
  normalize(aTE);
  cplx aTM[3] = {k[1] * aTE[2] - k[2] * aTE[1], k[2] * aTE[0] - k[0] * aTE[2],
                 k[0] * aTE[1] - k[1] * aTE[0]};  // a_TM = k_inc x a_TE (rcwa.cpp:491)
  normalize(aTM);
  cplx pol[3];
  for (int i = 0; i < 3; ++i) pol[i] = pte * aTE[i] + ptm * aTM[i];
  normalize(pol);                                    // rcwa.cpp:495-498
  // csrc = W_ref^-1 esrc = esrc (W_ref = I): entries PQ/2 and PQ/2 + PQ.
  const int n = 2 * PQ;
  std::vector<cplx> r(3 * PQ), t(3 * PQ);
  for (int i = 0; i < n; ++i) {
    r[i] = s11c[i] * pol[0] + s11c[n + i] * pol[1];  // rcwa.cpp:535
    t[i] = s21c[i] * pol[0] + s21c[n + i] * pol[1];  // rcwa.cpp:553
  }
  for (int i = 0; i < PQ; ++i) {
    r[2 * PQ + i] = -(wv.Kx[i] * r[i] + wv.Ky[i] * r[PQ + i]) / wv.Kz_ref[i];  // rcwa.cpp:541
    t[2 * PQ + i] = -(wv.Kx[i] * t[i] + wv.Ky[i] * t[PQ + i]) / wv.Kz_trn[i];  // rcwa.cpp:560
  }
  Efficiencies e;
  e.R.resize(PQ);
  e.T.resize(PQ);
  const double den = (k[2] / m.ur_ref).real();
  for (int i = 0; i < PQ; ++i) {                     // rcwa.cpp:574-583
    const double rr = std::norm(r[i]) + std::norm(r[PQ + i]) + std::norm(r[2 * PQ + i]);
    const double tt = std::norm(t[i]) + std::norm(t[PQ + i]) + std::norm(t[2 * PQ + i]);
    e.R[i] = (-wv.Kz_ref[i] / m.ur_ref).real() / den * rr;
    e.T[i] = (wv.Kz_trn[i] / m.ur_trn).real() / den * tt;
    e.R_tot += e.R[i];
    e.T_tot += e.T[i];
  }
  e.r = std::move(r);
  e.t = std::move(t);
  return e;
}

}  // namespace rcwa_host
