// Tangent-linear kernels: the derivative of the hot path along a direction of its inputs (H(k), the DMFT
// self-energy, mu, U and U~, chi_loc), for the gradient of chi(q, omega) and Sigma_DGA(k, nu) with respect
// to model parameters.  Each derivative is the product rule applied to one step of the forward path:
//   G^-1 = (i nu + mu) - H - Sigma                 dG = G (dH + dSigma - dmu) G
//   w = -beta g_i g_j, winv = 1/w                  dw = -beta (dg_i g_j + g_i dg_j), dwinv = -dw / w^2
//   C = sym(chi_loc) diag(winv) - 1               dC = sym(dchi_loc) diag(winv) + sym(chi_loc) diag(dwinv)
//   gamma = sum_nu C[nu rows], 1 + gamma           dgamma = sum_nu dC[nu rows]
//   chi0q = -beta/nkp sum_k G (x) G(k - q)         dchi0q = bubble(dG, G) + bubble(G, dG)
//   S = chi0q diag(winv) - 1                       dS = dchi0q diag(winv) + chi0q diag(dwinv)
//   eta A = (1+gamma) S C,  A = 1 - S C            d eta A = dgamma S C + (1 + gamma + eta)(dS C + S dC)
// The last line is solved with the factors of A already formed (M = A^T: M d eta^T = right-hand side),
// so each direction costs a solve, not a factorisation.
#include "adga_kernels.h"
#include "small_inv.h"

namespace adga {

// dGt[dir][w][k][a*n + b] = (G X G)_ab, X = dH(k) + diag(dsiw(nu_w)) - dmu: the G table's derivative
template <int NN>
__global__ void __launch_bounds__(256) k_dgtable(Geom g, int ndir, const z64* __restrict__ gt,
                                                 const z64* __restrict__ dhk, const z64* __restrict__ dsiw,
                                                 const double* __restrict__ dmu, z64* __restrict__ dgt) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)g.nkp * g.nw;
  if (idx >= per * ndir) return;
  const int dir = (int)(idx / per);
  const long wk = idx % per;
  const int k = (int)(wk % g.nkp), w = (int)(wk / g.nkp);
  const int iw = g.glo + w;
  z64 G[NN][NN], X[NN][NN];
  const z64* gs = gt + wk * NN * NN;
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      G[i][j] = gs[i * NN + j];
      X[i][j] = dhk[((long)dir * g.nkp + k) * NN * NN + i + j * NN];  // dhk(i, j, k, dir), Fortran order
    }
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const z64 ds = dsiw[(iw + g.iwmax) + (long)i * 2 * g.iwmax + (long)dir * 2 * g.iwmax * NN];
    X[i][i] = cadd(X[i][i], mk(ds.re - dmu[dir], ds.im));
  }
  z64* out = dgt + ((long)dir * per + wk) * NN * NN;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    z64 t[NN];  // row i of G X
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      z64 a = mk(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < NN; ++m) a = cfma(G[i][m], X[m][j], a);
      t[j] = a;
    }
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      z64 a = mk(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < NN; ++m) a = cfma(t[m], G[m][j], a);
      out[i * NN + j] = a;
    }
  }
}

bool launch_dgtable(const Geom& g, int ndir, const z64* gt, const z64* dhk, const z64* dsiw, const double* dmu,
                    z64* dgt, hipStream_t s) {
  const long tot = (long)g.nkp * g.nw * ndir;
  const dim3 grid((unsigned)((tot + 255) / 256));
  switch (g.n) {
#define ADGA_DG(NN) \
  case NN: hipLaunchKernelGGL(k_dgtable<NN>, grid, dim3(256), 0, s, g, ndir, gt, dhk, dsiw, dmu, dgt); return true;
    ADGA_DG(1) ADGA_DG(2) ADGA_DG(3) ADGA_DG(4) ADGA_DG(5)
#undef ADGA_DG
    default: return false;
  }
}

// dw, dwinv (n2, nf) of one direction from giw and its derivative dgiw (2*iwmax, n)
__global__ void k_dlocal_bubble(Geom g, int iwb, const z64* __restrict__ giw, const z64* __restrict__ dgiw,
                                z64* __restrict__ dw, z64* __restrict__ dwinv) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.n2 * g.nf) return;
  const int s = idx % g.n2, d = idx / g.n2;
  const int i = s / g.n, j = s % g.n, iwf = d - g.iwf;
  const long ri = (iwf + g.iwmax) + (long)i * 2 * g.iwmax, rj = (iwf - iwb + g.iwmax) + (long)j * 2 * g.iwmax;
  const z64 gi = giw[ri], gj = giw[rj], dgi = dgiw[ri], dgj = dgiw[rj];
  const z64 w = cscale(cmul(gi, gj), -g.beta);
  const z64 dww = cscale(cadd(cmul(dgi, gj), cmul(gi, dgj)), -g.beta);
  dw[idx] = dww;
  const z64 r = crecip(w);
  dwinv[idx] = cscale(cmul(dww, cmul(r, r)), -1.0);
}

// dC(i1, i2) = sym(dX)(i1, i2) winv(i2) + sym(X)(i1, i2) dwinv(i2), Fortran (N, N); dX may be null (zero)
__global__ void k_dchi0wF(int N, const z64* __restrict__ x, const z64* __restrict__ dx, const z64* __restrict__ winv,
                          const z64* __restrict__ dwinv, z64* __restrict__ dc) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)N * N) return;
  const int i1 = (int)(idx % N), i2 = (int)(idx / N);
  const z64 a = x[(long)i1 + (long)i2 * N], b = x[(long)i2 + (long)i1 * N];
  z64 v = cmul(mk(0.5 * (a.re + b.re), 0.5 * (a.im + b.im)), dwinv[i2]);
  if (dx) {
    const z64 da = dx[(long)i1 + (long)i2 * N], db = dx[(long)i2 + (long)i1 * N];
    v = cfma(mk(0.5 * (da.re + db.re), 0.5 * (da.im + db.im)), winv[i2], v);
  }
  dc[idx] = v;
}

// dgamma(s, j) = sum_d dC(d*n2 + s, j), (n2, N) Fortran; d(1 + gamma) is the same array
__global__ void k_dgammaw(Geom g, const z64* __restrict__ dc, z64* __restrict__ dgam) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)g.n2 * g.N) return;
  const int s = (int)(idx % g.n2);
  const long j = idx / g.n2;
  z64 acc = mk(0.0, 0.0);
  for (int d = 0; d < g.nf; ++d) acc = cadd(acc, dc[(long)(d * g.n2 + s) + j * g.N]);
  dgam[idx] = acc;
}

// Derivatives of the local chi and the local bubble: beta^-2 (dgamma . chi0w + gamma . dchi0w) and
// beta^-2 sum_nu dchi0w on the diagonal.
__global__ void k_dlocal_chi(Geom g, const z64* __restrict__ gam_d, const z64* __restrict__ gam_m,
                             const z64* __restrict__ dgam_d, const z64* __restrict__ dgam_m,
                             const z64* __restrict__ w, const z64* __restrict__ dw, z64* __restrict__ out_d,
                             z64* __restrict__ out_m, z64* __restrict__ bubloc) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.n4) return;
  const int s1 = idx % g.n2, s2 = idx / g.n2;
  z64 ad = mk(0.0, 0.0), am = mk(0.0, 0.0), bl = mk(0.0, 0.0);
  for (int d = 0; d < g.nf; ++d) {
    const z64 wv = w[s2 + d * g.n2], dwv = dw[s2 + d * g.n2];
    const long col = (long)(d * g.n2 + s2) * g.n2;
    ad = cfma(dgam_d[s1 + col], wv, cfma(gam_d[s1 + col], dwv, ad));
    am = cfma(dgam_m[s1 + col], wv, cfma(gam_m[s1 + col], dwv, am));
    if (s1 == s2) bl = cadd(bl, dwv);
  }
  const double ib2 = 1.0 / (g.beta * g.beta);
  out_d[idx] = cscale(ad, ib2);
  out_m[idx] = cscale(am, ib2);
  bubloc[idx] = cscale(bl, ib2);
}

void launch_local_tangent(const Geom& g, int iwb, const z64* giw, const z64* dgiw, const z64* xd, const z64* xm,
                          const z64* dxd, const z64* dxm, const z64* winv, const z64* gamd, const z64* gamm,
                          const z64* chi0w, z64* dw, z64* dwinv, z64* dcd, z64* dcm, z64* dgamd, z64* dgamm,
                          z64* dchiloc_d, z64* dchiloc_m, z64* dbubloc, hipStream_t s) {
  hipLaunchKernelGGL(k_dlocal_bubble, dim3((g.n2 * g.nf + 255) / 256), dim3(256), 0, s, g, iwb, giw, dgiw, dw, dwinv);
  const long nn = (long)g.N * g.N;
  hipLaunchKernelGGL(k_dchi0wF, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, g.N, xd, dxd, winv, dwinv, dcd);
  hipLaunchKernelGGL(k_dchi0wF, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, g.N, xm, dxm, winv, dwinv, dcm);
  const long pm = (long)g.n2 * g.N;
  hipLaunchKernelGGL(k_dgammaw, dim3((unsigned)((pm + 255) / 256)), dim3(256), 0, s, g, dcd, dgamd);
  hipLaunchKernelGGL(k_dgammaw, dim3((unsigned)((pm + 255) / 256)), dim3(256), 0, s, g, dcm, dgamm);
  hipLaunchKernelGGL(k_dlocal_chi, dim3((g.n4 + 255) / 256), dim3(256), 0, s, g, gamd, gamm, dgamd, dgamm, chi0w, dw,
                     dchiloc_d, dchiloc_m, dbubloc);
}

// dS[b][d][a][c] = dchi0q(a, c, d) winv(c, d) + chi0q(a, c, d) dwinv(c, d)
__global__ void k_dsmat(Geom g, int npts, const z64* __restrict__ chi0q, const z64* __restrict__ dchi0q,
                        const z64* __restrict__ winv, const z64* __restrict__ dwinv, z64* __restrict__ dS) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)npts * g.nf * g.n4) return;
  const int e = (int)(idx % g.n4);
  const long bd = idx / g.n4;
  const int d = (int)(bd % g.nf);
  const int a = e / g.n2, c = e % g.n2;
  const long src = bd * g.n4 + a + (long)c * g.n2;
  dS[idx] = cfma(chi0q[src], dwinv[c + d * g.n2], cmul(dchi0q[src], winv[c + d * g.n2]));
}

void launch_dsmat(const Geom& g, int npts, const z64* chi0q, const z64* dchi0q, const z64* winv, const z64* dwinv,
                  z64* dS, hipStream_t s) {
  const long tot = (long)npts * g.nf * g.n4;
  hipLaunchKernelGGL(k_dsmat, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, chi0q, dchi0q, winv, dwinv,
                     dS);
}

// Right operands of the derivative's right-hand side, in the residual product's layout
// Z[(ch N + x) ncol + b n2 + s] (x = d n2 + i'):
//   Z1 = S^T dopg^T + dS^T (opg + eta)^T,  Z2 = S^T (opg + eta)^T,  so that  rhs = C^T Z1 + dC^T Z2,
// with (opg + eta)^T = opg^T + Y (Y: the solution, eta^T).
__global__ void k_tan_z(Geom g, int npts, const z64* __restrict__ S, const z64* __restrict__ dS,
                        const z64* __restrict__ opg_d, const z64* __restrict__ opg_m, const z64* __restrict__ dopg_d,
                        const z64* __restrict__ dopg_m, const z64* __restrict__ Y, z64* __restrict__ Z1,
                        z64* __restrict__ Z2) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ncol = (long)npts * g.n2;
  if (idx >= 2 * g.N * ncol) return;
  const int s = (int)(idx % g.n2);
  const long mx = idx / g.n2;
  const int x = (int)(mx % g.N), m = (int)(mx / g.N);
  const int b = m / 2, ch = m % 2;
  const int d = x / g.n2, ip = x % g.n2;
  const z64* Sd = S + ((long)b * g.nf + d) * g.n4;
  const z64* dSd = dS + ((long)b * g.nf + d) * g.n4;
  const z64* opg = ch == 0 ? opg_d : opg_m;
  const z64* dopg = ch == 0 ? dopg_d : dopg_m;
  const z64* Ym = Y + (long)m * g.N * g.n2;
  z64 z1 = mk(0.0, 0.0), z2 = mk(0.0, 0.0);
  for (int cp = 0; cp < g.n2; ++cp) {
    const long col = (long)(d * g.n2 + cp);
    const z64 pe = cadd(opg[s + col * g.n2], Ym[col * g.n2 + s]);  // (opg + eta)(s, col)
    const z64 sv = Sd[cp * g.n2 + ip];
    z1 = cfma(sv, dopg[s + col * g.n2], z1);
    z1 = cfma(dSd[cp * g.n2 + ip], pe, z1);
    z2 = cfma(sv, pe, z2);
  }
  const long o = ((long)ch * g.N + x) * ncol + (long)b * g.n2 + s;
  Z1[o] = z1;
  Z2[o] = z2;
}

void launch_tangent_rhs(const Geom& g, int npts, const z64* S, const z64* dS, const z64* cd, const z64* cm,
                        const z64* dcd, const z64* dcm, const z64* opgd, const z64* opgm, const z64* dopgd,
                        const z64* dopgm, const z64* Y, z64* Z1, z64* Z2, z64* B, hipStream_t s) {
  const long tot = 2L * npts * g.N * g.n2;
  hipLaunchKernelGGL(k_tan_z, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, S, dS, opgd, opgm, dopgd,
                     dopgm, Y, Z1, Z2);
  resid_gemm(g, npts, 0, 2, cd, cm, Z1, B, nullptr, nullptr, s);
  resid_gemm(g, npts, 0, 2, dcd, dcm, Z2, B, nullptr, nullptr, s, true);
}

// Derivative of chi_ch(q) (non-local part) and of the non-local bubble:
//   beta^-2 [dgamma . (chi0q - chi0w) + gamma . (dchi0q - dchi0w) + deta . chi0q + eta . dchi0q],
//   beta^-2 (sum_nu dchi0q - sum_nu dchi0w)   (eta from Y^T, deta from dY^T).
__global__ void k_dchi(Geom g, int npts, const z64* __restrict__ chi0q, const z64* __restrict__ dchi0q,
                       const z64* __restrict__ chi0w, const z64* __restrict__ dchi0w, const z64* __restrict__ gam_d,
                       const z64* __restrict__ gam_m, const z64* __restrict__ dgam_d, const z64* __restrict__ dgam_m,
                       const z64* __restrict__ Y, const z64* __restrict__ dY, z64* __restrict__ dchi_out,
                       z64* __restrict__ dbub_out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)npts * 2 * g.n4) return;
  const int e = (int)(idx % g.n4);
  const int ch = (int)((idx / g.n4) % 2);
  const int b = (int)(idx / (2 * g.n4));
  const int s = e % g.n2, s2 = e / g.n2;
  const z64* gam = ch == 0 ? gam_d : gam_m;
  const z64* dgam = ch == 0 ? dgam_d : dgam_m;
  const long m = 2 * b + ch;
  const z64* Ym = Y + m * g.N * g.n2;
  const z64* dYm = dY + m * g.N * g.n2;
  const z64* cq = chi0q + (long)b * g.nf * g.n4;
  const z64* dcq = dchi0q + (long)b * g.nf * g.n4;
  z64 acc = mk(0.0, 0.0), bs = mk(0.0, 0.0);
  for (int d = 0; d < g.nf; ++d) {
    for (int i3 = 0; i3 < g.n2; ++i3) {
      const long x = (long)d * g.n2 + i3;
      const long ci = (long)d * g.n4 + i3 + s2 * g.n2;
      z64 c = cq[ci], dc = dcq[ci];
      z64 cnl = c, dcnl = dc;
      if (i3 == s2) {
        cnl = csub(c, chi0w[s2 + d * g.n2]);
        dcnl = csub(dc, dchi0w[s2 + d * g.n2]);
      }
      acc = cfma(dgam[s + x * g.n2], cnl, acc);
      acc = cfma(gam[s + x * g.n2], dcnl, acc);
      acc = cfma(dYm[x * g.n2 + s], c, acc);
      acc = cfma(Ym[x * g.n2 + s], dc, acc);
    }
    if (ch == 0) {
      bs = cadd(bs, dcq[(long)d * g.n4 + s + s2 * g.n2]);
      if (s == s2) bs = csub(bs, dchi0w[s2 + d * g.n2]);
    }
  }
  const double ib2 = 1.0 / (g.beta * g.beta);
  dchi_out[((long)b * 2 + ch) * g.n4 + e] = cscale(acc, ib2);
  if (ch == 0) dbub_out[(long)b * g.n4 + e] = cscale(bs, ib2);
}

void launch_dchi(const Geom& g, int npts, const z64* chi0q, const z64* dchi0q, const z64* chi0w, const z64* dchi0w,
                 const z64* gamd, const z64* gamm, const z64* dgamd, const z64* dgamm, const z64* Y, const z64* dY,
                 z64* dchi_out, z64* dbub_out, hipStream_t s) {
  const long tot = (long)npts * 2 * g.n4;
  hipLaunchKernelGGL(k_dchi, dim3((unsigned)((tot + 127) / 128)), dim3(128), 0, s, g, npts, chi0q, dchi0q, chi0w,
                     dchi0w, gamd, gamm, dgamd, dgamm, Y, dY, dchi_out, dbub_out);
}

}  // namespace adga
