// k-grid convolutions by separable DFT: the bubble for every q at once and the EOM sum over q.
//
// Both are convolutions on the periodic k grid (k - q folds back by integer modulo,
// kq_tools.f90:96-108):
//   bubble  c(q) = sum_k a(k) b(k - q)    a = G(., nu)_il,  b = G(., nu - omega)_k'j
//                = 1/nkp sum_R A(-R) B(R) e^{-iqR}
//   EOM     s(k) = sum_q f(q) g(k - q)    f = m_tot(q)_{...},  g = G(., nu - omega)_jk'
//                = 1/nkp sum_R F(R) G(R) e^{+ikR}
// with the forward transform X(R) = sum_k x(k) e^{-ikR}, e^{-ikR} = prod_axes e^{-2 pi i k_a R_a / n_a}.
// The transform is applied one axis at a time (a 3D DFT is three batched products with the
// n_a x n_a twiddle matrix); each pass is bound by HBM streaming of the batch of fields.
// A field is one contiguous grid (kz fastest, like k_index_3, kq_tools.f90:119-125).
#include "adga_kernels.h"

namespace adga {

// out[f][x][y][z] = sum_k' W_axis[R][k'] in[f][..k'..] with R the output coordinate on `axis`.
// tw holds the forward twiddles e^{-2 pi i R k / n} of the three axes back to back; sign > 0 uses
// their conjugates.  One thread per output element.
__global__ void __launch_bounds__(256) k_dft_axis(long nfield, int nx, int ny, int nz, int axis, int sign,
                                                  const z64* __restrict__ tw, const z64* __restrict__ in,
                                                  z64* __restrict__ out) {
  const long per = (long)nx * ny * nz;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nfield * per) return;
  long f = idx / per;
  int e = (int)(idx % per);
  int z = e % nz, y = (e / nz) % ny, x = e / (ny * nz);
  const z64* src = in + f * per;
  int n, R, stride;
  long base;
  const z64* w;
  if (axis == 2) {
    n = nz; R = z; stride = 1; base = (long)x * ny * nz + (long)y * nz; w = tw + nx * nx + ny * ny;
  } else if (axis == 1) {
    n = ny; R = y; stride = nz; base = (long)x * ny * nz + z; w = tw + nx * nx;
  } else {
    n = nx; R = x; stride = ny * nz; base = (long)y * nz + z; w = tw;
  }
  const z64* wr = w + (long)R * n;
  z64 acc = mk(0.0, 0.0);
  if (sign < 0) {
    for (int k = 0; k < n; ++k) acc = cfma(wr[k], src[base + (long)k * stride], acc);
  } else {
    for (int k = 0; k < n; ++k) acc = cfma(cconj(wr[k]), src[base + (long)k * stride], acc);
  }
  out[idx] = acc;
}

// The same with one thread per line (the NK outputs along `axis` of one field and one pair of the other
// coordinates): the line's inputs are read once into registers instead of once per output, and consecutive
// lanes take consecutive lines along the innermost of the other coordinates, so the passes along x and y
// load and store whole rows.  Each output is the same sum, in the same order, as in k_dft_axis.
template <int NK>
__global__ void __launch_bounds__(256) k_dft_line(long nfield, int nx, int ny, int nz, int axis, int sign,
                                                  const z64* __restrict__ tw, const z64* __restrict__ in,
                                                  z64* __restrict__ out) {
  const long per = (long)nx * ny * nz;
  long lid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (lid >= nfield * (per / NK)) return;
  long base;
  int stride;
  const z64* w;
  if (axis == 2) {  // lines along z (unit stride), lanes over y then x
    const int y = (int)(lid % ny);
    lid /= ny;
    const int x = (int)(lid % nx);
    base = (lid / nx) * per + (long)x * ny * nz + (long)y * nz;
    stride = 1;
    w = tw + nx * nx + ny * ny;
  } else if (axis == 1) {  // along y, lanes over z then x
    const int z = (int)(lid % nz);
    lid /= nz;
    const int x = (int)(lid % nx);
    base = (lid / nx) * per + (long)x * ny * nz + z;
    stride = nz;
    w = tw + nx * nx;
  } else {  // along x, lanes over z then y
    const int z = (int)(lid % nz);
    lid /= nz;
    const int y = (int)(lid % ny);
    base = (lid / ny) * per + (long)y * nz + z;
    stride = ny * nz;
    w = tw;
  }
  z64 v[NK];
#pragma unroll
  for (int k = 0; k < NK; ++k) v[k] = in[base + (long)k * stride];
#pragma unroll
  for (int R = 0; R < NK; ++R) {
    z64 acc = mk(0.0, 0.0);
#pragma unroll
    for (int k = 0; k < NK; ++k) acc = cfma(sign < 0 ? w[R * NK + k] : cconj(w[R * NK + k]), v[k], acc);
    out[base + (long)R * stride] = acc;
  }
}

static void dft_pass(long nfield, int nx, int ny, int nz, int axis, int sign, const z64* tw, const z64* in, z64* out,
                     hipStream_t s) {
  const int n = axis == 2 ? nz : axis == 1 ? ny : nx;
  const long nline = nfield * ((long)nx * ny * nz / n);
  const dim3 lg((unsigned)((nline + 255) / 256));
  switch (n) {
#define ADGA_DL(NK)                                                                                                \
  case NK:                                                                                                        \
    hipLaunchKernelGGL(k_dft_line<NK>, lg, dim3(256), 0, s, nfield, nx, ny, nz, axis, sign, tw, in, out);        \
    return;
    ADGA_DL(2) ADGA_DL(3) ADGA_DL(4) ADGA_DL(5) ADGA_DL(6) ADGA_DL(7) ADGA_DL(8) ADGA_DL(9) ADGA_DL(10)
    ADGA_DL(11) ADGA_DL(12) ADGA_DL(13) ADGA_DL(14) ADGA_DL(15) ADGA_DL(16) ADGA_DL(17) ADGA_DL(18)
    ADGA_DL(19) ADGA_DL(20) ADGA_DL(21) ADGA_DL(22) ADGA_DL(23) ADGA_DL(24)
#undef ADGA_DL
    default:
      break;
  }
  const long tot = nfield * nx * ny * nz;
  hipLaunchKernelGGL(k_dft_axis, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, nfield, nx, ny, nz, axis, sign,
                     tw, in, out);
}

// Three axis passes: a -> b -> a -> b; the transform ends in b.
void dft3(const Geom& g, long nfield, int sign, const z64* tw, z64* a, z64* b, hipStream_t s) {
  dft_pass(nfield, g.nkx, g.nky, g.nkz, 2, sign, tw, a, b, s);
  dft_pass(nfield, g.nkx, g.nky, g.nkz, 1, sign, tw, b, a, s);
  dft_pass(nfield, g.nkx, g.nky, g.nkz, 0, sign, tw, a, b, s);
}

// G table [w][k][ab] -> fields [w][ab][k]
__global__ void k_gt_fields(Geom g, const z64* __restrict__ gt, z64* __restrict__ out) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)g.nw * g.n2 * g.nkp;
  if (idx >= tot) return;
  int k = (int)(idx % g.nkp);
  long wab = idx / g.nkp;
  int ab = (int)(wab % g.n2);
  long w = wab / g.n2;
  out[idx] = gt[(w * g.nkp + k) * g.n2 + ab];
}

// Ghat[w][ab][R] = sum_k G(k, nu_w)_ab e^{-ikR}, for every frequency of the G table.
void launch_ghat(const Geom& g, const z64* tw, const z64* gt, z64* tmp, z64* ghat, hipStream_t s) {
  long tot = (long)g.nw * g.n2 * g.nkp;
  hipLaunchKernelGGL(k_gt_fields, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, gt, tmp);
  dft3(g, (long)g.nw * g.n2, -1, tw, tmp, ghat, s);
}

// ---------------------------------------------------------------- bubble for every q
// P[(d, c)][R] = Ghat(nu_d)_il(-R) Ghat(nu_d - omega)_k'j(R), c = ((i*n + l)*n + k')*n + j
__global__ void k_bub_prod(Geom g, int iwb, const z64* __restrict__ ghat, z64* __restrict__ P) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)g.nf * g.n4 * g.nkp;
  if (idx >= tot) return;
  int R = (int)(idx % g.nkp);
  long dc = idx / g.nkp;
  int c = (int)(dc % g.n4);
  int d = (int)(dc / g.n4);
  int il = c / g.n2, kj = c % g.n2;
  int rz = R % g.nkz, ry = (R / g.nkz) % g.nky, rx = R / (g.nky * g.nkz);
  int mR = ((g.nkz - rz) % g.nkz) + ((g.nky - ry) % g.nky) * g.nkz + ((g.nkx - rx) % g.nkx) * g.nky * g.nkz;
  int iwf = d - g.iwf;
  const z64* a = ghat + ((long)(iwf - g.glo) * g.n2 + il) * g.nkp;
  const z64* b = ghat + ((long)(iwf - iwb - g.glo) * g.n2 + kj) * g.nkp;
  P[idx] = cmul(a[mR], b[R]);
}

void launch_bubble_all(const Geom& g, int iwb, const z64* tw, const z64* ghat, z64* P, z64* out, hipStream_t s) {
  long tot = (long)g.nf * g.n4 * g.nkp;
  hipLaunchKernelGGL(k_bub_prod, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, iwb, ghat, P);
  dft3(g, (long)g.nf * g.n4, -1, tw, P, out, s);
}

// the bubble's derivative by the product rule, for every q:
// P[(d, c)][R] = Ghat(nu_d)_il(-R) dGhat(nu_d - omega)_k'j(R) + dGhat(nu_d)_il(-R) Ghat(nu_d - omega)_k'j(R)
__global__ void k_bub_prod_tangent(Geom g, int iwb, const z64* __restrict__ ghat, const z64* __restrict__ dghat,
                                   z64* __restrict__ P) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)g.nf * g.n4 * g.nkp;
  if (idx >= tot) return;
  int R = (int)(idx % g.nkp);
  long dc = idx / g.nkp;
  int c = (int)(dc % g.n4);
  int d = (int)(dc / g.n4);
  int il = c / g.n2, kj = c % g.n2;
  int rz = R % g.nkz, ry = (R / g.nkz) % g.nky, rx = R / (g.nky * g.nkz);
  int mR = ((g.nkz - rz) % g.nkz) + ((g.nky - ry) % g.nky) * g.nkz + ((g.nkx - rx) % g.nkx) * g.nky * g.nkz;
  int iwf = d - g.iwf;
  const long oa = ((long)(iwf - g.glo) * g.n2 + il) * g.nkp, ob = ((long)(iwf - iwb - g.glo) * g.n2 + kj) * g.nkp;
  P[idx] = cadd(cmul(ghat[oa + mR], dghat[ob + R]), cmul(dghat[oa + mR], ghat[ob + R]));
}

void launch_bubble_all_tangent(const Geom& g, int iwb, const z64* tw, const z64* ghat, const z64* dghat, z64* P,
                               z64* out, hipStream_t s) {
  long tot = (long)g.nf * g.n4 * g.nkp;
  hipLaunchKernelGGL(k_bub_prod_tangent, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, iwb, ghat, dghat, P);
  dft3(g, (long)g.nf * g.n4, -1, tw, P, out, s);
}

// chi0q[b][d](i1 = i*n + j, i2 = l*n + k') = -beta/nkp^2 all[(d, c)][q_b]
// (the -beta/nkp of accumulate_chi0 and main.F90:653, and the 1/nkp of the convolution theorem)
__global__ void k_bub_gather(Geom g, int npts, const int* __restrict__ qidx, const z64* __restrict__ all,
                             z64* __restrict__ chi0q) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)npts * g.nf * g.n4;
  if (idx >= tot) return;
  int c = (int)(idx % g.n4);
  long bd = idx / g.n4;
  int d = (int)(bd % g.nf);
  int b = (int)(bd / g.nf);
  int i = c / (g.n * g.n2), l = (c / g.n2) % g.n, kp = (c / g.n) % g.n, j = c % g.n;
  double f = -g.beta / ((double)g.nkp * g.nkp);
  z64 v = all[((long)d * g.n4 + c) * g.nkp + qidx[b]];
  chi0q[bd * g.n4 + (i * g.n + j) + (long)(l * g.n + kp) * g.n2] = cscale(v, f);
}

void launch_bubble_gather(const Geom& g, int npts, const int* qidx, const z64* all, z64* chi0q, hipStream_t s) {
  long tot = (long)npts * g.nf * g.n4;
  hipLaunchKernelGGL(k_bub_gather, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, npts, qidx, all, chi0q);
}

// ---------------------------------------------------------------- EOM over q
// Macc[(nu, i1, l, k')][q_b] += m_b(i1, n2*(iwf + nu) + l*n + k') for nu = 0..iwf-1
// (m_tot of eom_module.f90:187-201 placed on the q grid; points of a batch are summed in order).
__global__ void k_eom_scatter(Geom g, int npts, const int* __restrict__ qidx, const z64* __restrict__ mt,
                              z64* __restrict__ macc) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long nfield = (long)g.iwf * g.n4;
  if (idx >= nfield) return;
  int kp = (int)(idx % g.n), l = (int)((idx / g.n) % g.n);
  int i1 = (int)((idx / g.n2) % g.n2);
  int nu = (int)(idx / g.n4);
  long col = (long)g.n2 * (g.iwf + nu) + l * g.n + kp;
  long per = (long)g.n2 * g.N;
  z64* dst = macc + idx * g.nkp;
  for (int b = 0; b < npts; ++b) {
    int q = qidx[b];
    dst[q] = cadd(dst[q], mt[(long)b * per + i1 + col * g.n2]);
  }
}

// The same when the batch's points are distinct (each q at most once): a workgroup takes pt points at one
// nu, reads their n^4 values of m_tot at that nu (contiguous per point, index i1 + n2 (l n + k')) into LDS
// with coalesced loads, and adds them to Macc with consecutive points on consecutive lanes, so that the
// consecutive q of a usual batch give whole-line accesses; every element of Macc gets the same single
// addition as in k_eom_scatter.
__global__ void __launch_bounds__(256) k_eom_scatter_t(Geom g, int npts, int pt, const int* __restrict__ qidx,
                                                       const z64* __restrict__ mt, z64* __restrict__ macc) {
  extern __shared__ z64 tile[];  // [pt][n4 + 1]
  const int nu = blockIdx.y, b0 = blockIdx.x * pt, t = threadIdx.x;
  const int n2 = g.n2, n4 = g.n4;
  const int np = min(pt, npts - b0);
  const long per = (long)n2 * g.N;
  const long off = (long)n4 * (g.iwf + nu);  // the (point, nu) block of m_tot
  for (int e = t; e < np * n4; e += blockDim.x) {
    const int pb = e / n4, f = e % n4;
    tile[pb * (n4 + 1) + f] = mt[(long)(b0 + pb) * per + off + f];
  }
  __syncthreads();
  for (int e = t; e < n4 * pt; e += blockDim.x) {
    const int pb = e % pt, fm = e / pt;  // fm = i1 n2 + l n + k', the field within nu
    if (pb >= np) continue;
    const int i1 = fm / n2, lk = fm % n2;
    z64* dst = macc + ((long)nu * n4 + fm) * g.nkp + qidx[b0 + pb];
    *dst = cadd(*dst, tile[pb * (n4 + 1) + lk * n2 + i1]);
  }
}

void launch_eom_scatter(const Geom& g, int npts, const int* qidx, const z64* mt, z64* macc, hipStream_t s,
                        bool distinct) {
  if (distinct) {
    const int pt = std::max(1, std::min(32, 2688 / (g.n4 + 1)));  // <= 42 KB of LDS
    const size_t lds = (size_t)pt * (g.n4 + 1) * sizeof(z64);
    hipLaunchKernelGGL(k_eom_scatter_t, dim3((unsigned)((npts + pt - 1) / pt), (unsigned)g.iwf), dim3(256), lds, s, g,
                       npts, pt, qidx, mt, macc);
    return;
  }
  long nfield = (long)g.iwf * g.n4;
  hipLaunchKernelGGL(k_eom_scatter, dim3((unsigned)((nfield + 63) / 64)), dim3(64), 0, s, g, npts, qidx, mt, macc);
}

// H[(nu, i, l)][R] = sum_{j,k'} Mhat[(nu, i*n + j, l, k')][R] Ghat(nu - omega)_jk'(R)
__global__ void k_eom_contract(Geom g, int iwb, const z64* __restrict__ mhat, const z64* __restrict__ ghat,
                               z64* __restrict__ H, int accumulate) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)g.iwf * g.n2 * g.nkp;
  if (idx >= tot) return;
  int R = (int)(idx % g.nkp);
  long f = idx / g.nkp;
  int il = (int)(f % g.n2);
  int nu = (int)(f / g.n2);
  int i = il / g.n, l = il % g.n;
  const z64* gw = ghat + (long)(nu - iwb - g.glo) * g.n2 * g.nkp;
  z64 acc = mk(0.0, 0.0);
  for (int j = 0; j < g.n; ++j)
    for (int kp = 0; kp < g.n; ++kp) {
      long mf = (((long)nu * g.n2 + i * g.n + j) * g.n + l) * g.n + kp;
      acc = cfma(mhat[mf * g.nkp + R], gw[(long)(j * g.n + kp) * g.nkp + R], acc);
    }
  H[idx] = accumulate ? cadd(H[idx], acc) : acc;
}

// Sigma_nl(i, l, iwf + nu, kk) += 1/nkp Hk[(nu, i, l)][k_eom(kk)]
__global__ void k_eom_add(Geom g, const int* __restrict__ keom, const z64* __restrict__ hk, z64* __restrict__ sigma) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)g.n2 * g.iwf * g.nkp_eom;
  if (idx >= tot) return;
  int il = (int)(idx % g.n2);  // i + l*n (Fortran order of sigma)
  int nu = (int)((idx / g.n2) % g.iwf);
  long kk = idx / ((long)g.n2 * g.iwf);
  int i = il % g.n, l = il / g.n;
  int k = keom[kk] - 1;
  z64 v = cscale(hk[((long)nu * g.n2 + i * g.n + l) * g.nkp + k], 1.0 / g.nkp);
  long o = i + (long)l * g.n + (long)(g.iws + nu) * g.n2 + kk * g.n2 * (2 * g.iws);
  sigma[o] = cadd(sigma[o], v);
}

// Flush of one omega: transform the accumulated m_tot, contract with Ghat, transform back, add.
void launch_eom_flush(const Geom& g, int iwb, const int* keom, const z64* tw, const z64* ghat, z64* macc, z64* work,
                      z64* sigma, hipStream_t s, const EomTangent* tan) {
  long nfield = (long)g.iwf * g.n4;
  dft3(g, nfield, -1, tw, macc, work, s);  // Mhat in work (macc is scratch afterwards)
  long th = (long)g.iwf * g.n2 * g.nkp;
  long ta = (long)g.n2 * g.iwf * g.nkp_eom;
  const dim3 gth((unsigned)((th + 255) / 256)), gta((unsigned)((ta + 255) / 256));
  // derivatives along each tangent direction d: d Sigma = conv(dm_d, G) + conv(m, dG_d), i.e. in real
  // space dMhat_d Ghat + Mhat dGhat_d (each direction's accumulator is scratch afterwards)
  for (int d = 0; tan && d < tan->ndir; ++d) {
    z64* dm = tan->macc + d * tan->nm;
    dft3(g, nfield, -1, tw, dm, tan->work, s);  // dMhat_d in tan->work
    hipLaunchKernelGGL(k_eom_contract, gth, dim3(256), 0, s, g, iwb, tan->work, ghat, dm, 0);
    hipLaunchKernelGGL(k_eom_contract, gth, dim3(256), 0, s, g, iwb, work, tan->dghat + d * tan->ng, dm, 1);
    dft3(g, (long)g.iwf * g.n2, +1, tw, dm, tan->work, s);
    hipLaunchKernelGGL(k_eom_add, gta, dim3(256), 0, s, g, keom, tan->work, tan->sigma + d * tan->nsig);
  }
  hipLaunchKernelGGL(k_eom_contract, gth, dim3(256), 0, s, g, iwb, work, ghat, macc, 0);
  dft3(g, (long)g.iwf * g.n2, +1, tw, macc, work, s);  // H(k) in work
  hipLaunchKernelGGL(k_eom_add, gta, dim3(256), 0, s, g, keom, work, sigma);
}

}  // namespace adga
