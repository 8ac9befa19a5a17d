// Multi-determinant overlaps and Green's function by ratio updates relative
// to determinant 0 (the role of OmegaQMC/observables/wick.py:239-284,
// 425-485 for the step of afqmc_gto.py:842-935; same results as the explicit
// per-det path of greens_th4.py:75-148, 186-228).
//
// Det I differs from det 0 by k <= 4 orbitals: holes at positions R of occ_0
// replaced by particles P.  With A_0 = phi[occ_0, :] (S_0 = A_0^T) and X =
// A_0^-1, the matrix determinant lemma gives
//   det A_I' = det A_0 det Q_I,   Q_I[m][n] = (phi[p_m] X)[r_n],
// where A_I' is A_0 with rows R replaced in place (det S_I = sign_I det A_I'
// for the permutation that sorts A_I' rows into occ_I's order), and
// Woodbury gives the inverse as a rank-k correction
//   A_I'^-1 = X - X[:, R] Q_I^-1 F_I,   F_I = phi[P] X - E_R^T.
// The multi-det G needs only B = sum_I w_I D_I' A_I'^-T / overlap (G = B
// phi^T), so each kept det adds w_I X^T plus the rank-k term to the rows of B
// it occupies; w_I Q_I^-1 = c_I^* o_I,other sign_I det S_0 adj(Q_I) is formed
// from the adjugate, with no division by det Q_I.  One Gauss-Jordan inverse
// per walker and spin (the register elimination of the single-det kernel)
// replaces one per (walker, det) pair.
#include <algorithm>
#include <set>
#include <vector>

#include "afq_gemm.h"
#include "afq_gj.h"
#include "afq_kernels.h"
#include "afq_multidet.h"

namespace afq {

namespace {

constexpr int KMAX = 4;
constexpr int EXC = 2 + 2 * KMAX;  // k, holes r[4], particles p[4], sign

struct RatioArgs {
  int mode;  // 0: overlaps; 1: also B_s
  int W, M, nd, chunk;
  int N[2];
  const cplx* phi[2];
  const int* occ0[2];   // occ of det 0 (N_s ints)
  const int* exc[2];    // [nd][EXC]
  const int* slot[2];   // [M]: row slot of orbital p in the compact B, -1 if never occupied
  int nslot[2];
  const cplx* ci;
  cplx* ovlp;           // (W,)
  cplx* B[2];           // (W, M, N_s)
  // LDS offsets (bytes)
  int off_od, off_z, off_keep, off_ov;
  int off_order[2], off_inv[2], off_q[2], off_bc[2];
  int ld;               // row pitch of the inverse in LDS
};

__device__ __forceinline__ cplx det2(cplx a, cplx b, cplx c, cplx d) {
  return csub(cmul(a, d), cmul(b, c));
}

// determinant of the k x k matrix Q (row-major, k <= 4) without row `sr` and
// column `sc` (sr, sc = -1: none)
__device__ cplx small_det(const cplx* Q, int k, int sr, int sc) {
  cplx a[16];
  int n = 0;
  for (int i = 0; i < k; ++i) {
    if (i == sr) continue;
    for (int j = 0; j < k; ++j) {
      if (j == sc) continue;
      a[n++] = Q[i * k + j];
    }
  }
  const int m = k - (sr >= 0 ? 1 : 0);
  if (m == 0) return mkc(1.0, 0.0);
  if (m == 1) return a[0];
  if (m == 2) return det2(a[0], a[1], a[2], a[3]);
  if (m == 3) {
    cplx d = cmul(a[0], det2(a[4], a[5], a[7], a[8]));
    d = csub(d, cmul(a[1], det2(a[3], a[5], a[6], a[8])));
    return cadd(d, cmul(a[2], det2(a[3], a[4], a[6], a[7])));
  }
  // m == 4: expansion along row 0 over 3 x 3 minors
  cplx d = mkc(0.0, 0.0);
  for (int j = 0; j < 4; ++j) {
    cplx b[9];
    int t = 0;
    for (int i = 1; i < 4; ++i)
      for (int c = 0; c < 4; ++c)
        if (c != j) b[t++] = a[i * 4 + c];
    cplx m3 = cmul(b[0], det2(b[4], b[5], b[7], b[8]));
    m3 = csub(m3, cmul(b[1], det2(b[3], b[5], b[6], b[8])));
    m3 = cadd(m3, cmul(b[2], det2(b[3], b[4], b[6], b[7])));
    const cplx term = cmul(a[j], m3);
    d = (j & 1) ? csub(d, term) : cadd(d, term);
  }
  return d;
}

// Q[m][n] = sum_c phi[p_m][c] Inv[r_n][c] into Qs (k x k), every lane then
// reads it.  Lanes: four per (m, n) pair, each summing every fourth c.
__device__ void build_q(int k, const int* e, const cplx* phi, int N, const cplx* Inv, int ld,
                        cplx* Qs) {
  const int l = threadIdx.x & 63, pr = l >> 2, sub = l & 3;
  cplx acc = mkc(0.0, 0.0);
  if (pr < k * k) {
    const int m = pr / k, n = pr % k;
    const cplx* ph = phi + (long)e[1 + KMAX + m] * N;
    const cplx* iv = Inv + (long)e[1 + n] * ld;
    for (int c = sub; c < N; c += 4) cfma(acc, ph[c], iv[c]);
  }
  acc.x += __shfl_xor(acc.x, 1);
  acc.y += __shfl_xor(acc.y, 1);
  acc.x += __shfl_xor(acc.x, 2);
  acc.y += __shfl_xor(acc.y, 2);
  if (sub == 0 && pr < k * k) Qs[pr] = acc;
  gj::wave_sync();
}

template <int NB>
__global__ __launch_bounds__(128) void k_md_ratio(RatioArgs a) {
  extern __shared__ __align__(16) char smem[];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const long w = blockIdx.x;
  const int N = a.N[wv], nd = a.nd, ld = a.ld;
  cplx* od = (cplx*)(smem + a.off_od);     // [2][nd]
  cplx* zz = (cplx*)(smem + a.off_z);      // [nd]
  int* keep = (int*)(smem + a.off_keep);   // [nd]
  cplx* ovs = (cplx*)(smem + a.off_ov);    // [1]
  int* order = (int*)(smem + a.off_order[wv]);
  cplx* Inv = (cplx*)(smem + a.off_inv[wv]);
  cplx* Qs = (cplx*)(smem + a.off_q[wv]);  // [16] Q, [16] det S_0
  const cplx* phi = a.phi[wv] + w * a.M * N;
  const int* exc = a.exc[wv];
  if (N > 0) {
    cplx r[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j)
      r[j] = (l < N && j < N) ? phi[(long)a.occ0[wv][j] * N + l] : mkc(0.0, 0.0);
    int mystep = 0;
    gj::wave_gj<NB, false>(r, N, order, mystep, Qs + 16);
    gj::gj_store_inverse<NB>(r, N, mystep, order, Inv, ld);
  }
  // per-det overlap of this spin: sign_I det S_0 det Q_I
  const cplx det0 = N > 0 ? Qs[16] : mkc(1.0, 0.0);
  for (int I = 0; I < nd; ++I) {
    cplx o = mkc(1.0, 0.0);
    if (N > 0) {
      const int* e = exc + (long)I * EXC;
      const int k = e[0];
      build_q(k, e, phi, N, Inv, ld, Qs);
      const cplx dq = small_det(Qs, k, -1, -1);
      o = cmul(det0, dq);
      if (e[EXC - 1] < 0) o = mkc(-o.x, -o.y);
      gj::wave_sync();  // Qs is rewritten by the next det
    }
    if (l == 0) od[wv * nd + I] = o;
  }
  __syncthreads();
  // masks, weights and the overlap (greens_th4.py:27-46, 120-131), as the
  // explicit path: dets in chunks, chunk sums added in order
  if (wv == 0) {
    double om = 0.0;
    for (int I = l; I < nd; I += 64) {
      const cplx o = cmul(od[I], od[nd + I]);
      om = jmax(om, hypot(o.x, o.y));
    }
    for (int off = 32; off > 0; off >>= 1) om = jmax(om, __shfl_xor(om, off));
    const double tol = jmax(1e-14 * om, 1e-300);
    for (int I = l; I < nd; I += 64) {
      const cplx oa = od[I], ob = od[nd + I];
      const cplx o = cmul(oa, ob);
      const cplx wi = cmul(cmul(cconj(a.ci[I]), oa), ob);
      const bool kp = hypot(o.x, o.y) >= tol;
      keep[I] = kp ? 1 : 0;
      zz[I] = kp ? wi : mkc(0.0, 0.0);
    }
    if (l == 0) {
      cplx acc = mkc(0.0, 0.0);
      for (int c0 = 0; c0 < nd; c0 += a.chunk) {
        cplx cs = mkc(0.0, 0.0);
        for (int I = c0; I < min(nd, c0 + a.chunk); ++I)
          cs = cadd(cs, cmul(cmul(cconj(a.ci[I]), od[I]), od[nd + I]));
        acc = cadd(acc, cs);
      }
      *ovs = acc;
      a.ovlp[w] = acc;
    }
  }
  __syncthreads();
  if (a.mode == 0 || N == 0) return;  // no barrier after this point
  // B_s (compact rows) = sum over kept dets of w_I D_I' A_I'^-T
  const int ns = a.nslot[wv];
  cplx* Bc = (cplx*)(smem + a.off_bc[wv]);  // [ns][N]
  for (int e = l; e < ns * N; e += 64) Bc[e] = mkc(0.0, 0.0);
  gj::wave_sync();
  const cplx* oth = od + (1 - wv) * nd;
  for (int I = 0; I < nd; ++I) {
    if (!keep[I]) continue;  // uniform
    const int* e = exc + (long)I * EXC;
    const int k = e[0];
    const cplx zi = zz[I];
    // beta adj(Q_I) = w_I Q_I^-1 with beta = c_I^* o_I,other sign_I det S_0
    cplx beta = cmul(cmul(cconj(a.ci[I]), oth[I]), det0);
    if (e[EXC - 1] < 0) beta = mkc(-beta.x, -beta.y);
    cplx wq[16];
    if (k > 0) {
      build_q(k, e, phi, N, Inv, ld, Qs);
      for (int n = 0; n < k; ++n)
        for (int m = 0; m < k; ++m) {
          // adj(Q)[n][m] = (-1)^(n+m) det(Q without row m, column n)
          cplx c = small_det(Qs, k, m, n);
          if ((n + m) & 1) c = mkc(-c.x, -c.y);
          wq[n * KMAX + m] = cmul(beta, c);
        }
    }
    if (l < N) {
      // F[m][j] = (phi[p_m] X)[j] - [j == r_m], X[c][j] = Inv[j][c]; t = (w Q^-1 F)[:, j]
      cplx t[KMAX];
      int slot = l;  // row occ_0[j] sits in slot j
      for (int n = 0; n < KMAX; ++n) t[n] = mkc(0.0, 0.0);
      for (int m = 0; m < k; ++m) {
        const cplx* ph = phi + (long)e[1 + KMAX + m] * N;
        const cplx* iv = Inv + (long)l * ld;
        cplx v = mkc(0.0, 0.0);
        for (int c = 0; c < N; ++c) cfma(v, ph[c], iv[c]);
        if (l == e[1 + m]) {
          v.x -= 1.0;
          slot = a.slot[wv][e[1 + KMAX + m]];
        }
        for (int n = 0; n < k; ++n) cfma(t[n], wq[n * KMAX + m], v);
      }
      cplx* brow = Bc + (long)slot * N;
      const cplx* ivj = Inv + (long)l * ld;
      for (int c = 0; c < N; ++c) {
        cplx v = cmul(zi, ivj[c]);
        for (int n = 0; n < k; ++n) cfms(v, Inv[(long)e[1 + n] * ld + c], t[n]);
        brow[c] = cadd(brow[c], v);
      }
    }
    gj::wave_sync();  // Qs and the rows of Bc are touched again by the next det
  }
  const cplx ov = *ovs;
  cplx* Bg = a.B[wv] + w * a.M * N;
  const int* sl = a.slot[wv];
  for (long e = l; e < (long)a.M * N; e += 64) {
    const int p = (int)(e / N), c = (int)(e % N);
    const int s = sl[p];
    Bg[e] = s >= 0 ? cdiv(Bc[(long)s * N + c], ov) : mkc(0.0, 0.0);
  }
}

}  // namespace

struct MdRatio {
  int* exc[2] = {nullptr, nullptr};
  int* slot[2] = {nullptr, nullptr};
  int nslot[2] = {0, 0};
};

// Excitation tables of a multi-det system relative to det 0; returns false
// (and builds nothing) when some det is more than KMAX orbitals away.
int md_ratio_prepare(afq_system* sys, bool* ok) {
  *ok = false;
  hipStream_t s = sys->ctx->stream;
  const int M = sys->M, nd = sys->ndet;
  const int Ns[2] = {sys->Na, sys->Nb};
  const int* occd[2] = {sys->occ_a, sys->occ_b};
  std::vector<int> hexc[2], hslot[2];
  int nslot[2] = {0, 0};
  for (int sp = 0; sp < 2; ++sp) {
    const int N = Ns[sp];
    hexc[sp].assign((size_t)nd * EXC, 0);
    hslot[sp].assign(M, -1);
    if (N == 0) continue;
    std::vector<int> occ((size_t)nd * N);
    AFQ_CHECK_HIP(hipMemcpyAsync(occ.data(), occd[sp], sizeof(int) * occ.size(),
                                 hipMemcpyDeviceToHost, s));
    AFQ_CHECK_HIP(hipStreamSynchronize(s));
    const int* o0 = occ.data();
    for (int j = 0; j < N; ++j) hslot[sp][o0[j]] = j;
    int ns = N;
    const std::set<int> s0(o0, o0 + N);
    for (int I = 0; I < nd; ++I) {
      const int* oI = occ.data() + (size_t)I * N;
      const std::set<int> sI(oI, oI + N);
      std::vector<int> holes, parts;
      for (int j = 0; j < N; ++j)
        if (!sI.count(o0[j])) holes.push_back(j);
      for (int j = 0; j < N; ++j)
        if (!s0.count(oI[j])) parts.push_back(oI[j]);
      if ((int)holes.size() > KMAX) return AFQ_OK;  // ratio path not available
      const int k = (int)holes.size();
      std::vector<int> rep(o0, o0 + N);  // det 0 with the holes replaced in place
      for (int m = 0; m < k; ++m) rep[holes[m]] = parts[m];
      // sign of the permutation taking rep's order to occ_I's
      std::vector<int> pos(N);
      for (int j = 0; j < N; ++j)
        pos[j] = (int)(std::find(rep.begin(), rep.end(), oI[j]) - rep.begin());
      int inv = 0;
      for (int i = 0; i < N; ++i)
        for (int j = i + 1; j < N; ++j) inv += pos[i] > pos[j];
      int* e = hexc[sp].data() + (size_t)I * EXC;
      e[0] = k;
      for (int m = 0; m < k; ++m) {
        e[1 + m] = holes[m];
        e[1 + KMAX + m] = parts[m];
        if (hslot[sp][parts[m]] < 0) hslot[sp][parts[m]] = ns++;
      }
      e[EXC - 1] = (inv & 1) ? -1 : 1;
    }
    nslot[sp] = ns;
  }
  MdRatio* r = new MdRatio();
  for (int sp = 0; sp < 2; ++sp) {
    r->nslot[sp] = nslot[sp];
    if (hipMalloc(&r->exc[sp], sizeof(int) * std::max<size_t>(hexc[sp].size(), 1)) != hipSuccess ||
        hipMalloc(&r->slot[sp], sizeof(int) * M) != hipSuccess) {
      md_ratio_release(r);
      return fail(AFQ_ENOMEM, "multi-det excitation tables");
    }
    AFQ_CHECK_HIP(hipMemcpyAsync(r->exc[sp], hexc[sp].data(), sizeof(int) * hexc[sp].size(),
                                 hipMemcpyHostToDevice, s));
    AFQ_CHECK_HIP(hipMemcpyAsync(r->slot[sp], hslot[sp].data(), sizeof(int) * M,
                                 hipMemcpyHostToDevice, s));
  }
  AFQ_CHECK_HIP(hipStreamSynchronize(s));
  sys->md_ratio = r;
  *ok = true;
  return AFQ_OK;
}

void md_ratio_release(MdRatio* r) {
  if (!r) return;
  for (int sp = 0; sp < 2; ++sp) {
    if (r->exc[sp]) (void)hipFree(r->exc[sp]);
    if (r->slot[sp]) (void)hipFree(r->slot[sp]);
  }
  delete r;
}

// LDS layout of k_md_ratio (bytes), or 0 when it does not fit
static size_t ratio_lds(const afq_system* sys, int mode, int NB, RatioArgs* a) {
  const int nd = sys->ndet;
  size_t o = 0;
  auto take = [&](size_t n) {
    const size_t r = o;
    o += (n + 15) & ~size_t(15);
    return (int)r;
  };
  a->off_od = take(16UL * 2 * nd);
  a->off_z = take(16UL * nd);
  a->off_keep = take(4UL * nd);
  a->off_ov = take(16);
  a->ld = NB;
  const int Ns[2] = {sys->Na, sys->Nb};
  for (int sp = 0; sp < 2; ++sp) {
    a->off_order[sp] = take(256);
    a->off_inv[sp] = take(16UL * std::max(Ns[sp], 1) * NB);
    a->off_q[sp] = take(16UL * 32);
    a->off_bc[sp] = mode == 1 ? take(16UL * sys->md_ratio->nslot[sp] * Ns[sp]) : (int)o;
  }
  return o <= max_lds_bytes() ? o : 0;
}

bool md_ratio_fits(const afq_system* sys) {
  if (!sys->md_ratio) return false;
  RatioArgs a{};
  const int Nmax = std::max(sys->Na, sys->Nb);
  const int NB = Nmax <= 8 ? 8 : (Nmax + 3) / 4 * 4;
  return ratio_lds(sys, 1, NB, &a) > 0;
}

int md_ratio_launch(afq_system* sys, int mode, int W, const double* phia, const double* phib,
                    double* ovlp, double* Ba, double* Bb) {
  if (W <= 0) return AFQ_OK;
  RatioArgs a{};
  const int Nmax = std::max(sys->Na, sys->Nb);
  if (Nmax > gj::GJ_NMAX) return fail(AFQ_EUNSUPPORTED, "multi-det ratio: nocc %d", Nmax);
  const int NB = Nmax <= 8 ? 8 : (Nmax + 3) / 4 * 4;
  const size_t sm = ratio_lds(sys, mode, NB, &a);
  if (!sm) return fail(AFQ_EUNSUPPORTED, "multi-det ratio: LDS");
  a.mode = mode;
  a.W = W;
  a.M = sys->M;
  a.nd = sys->ndet;
  a.chunk = std::max(1, sys->det_chunk);
  a.N[0] = sys->Na;
  a.N[1] = sys->Nb;
  a.phi[0] = (const cplx*)phia;
  a.phi[1] = (const cplx*)phib;
  a.occ0[0] = sys->occ_a;
  a.occ0[1] = sys->occ_b;
  for (int sp = 0; sp < 2; ++sp) {
    a.exc[sp] = sys->md_ratio->exc[sp];
    a.slot[sp] = sys->md_ratio->slot[sp];
    a.nslot[sp] = sys->md_ratio->nslot[sp];
  }
  a.ci = (const cplx*)sys->ci;
  a.ovlp = (cplx*)ovlp;
  a.B[0] = (cplx*)Ba;
  a.B[1] = (cplx*)Bb;
  hipStream_t s = sys->ctx->stream;
#define AFQ_RATIO(NBV) \
  hipLaunchKernelGGL((k_md_ratio<NBV>), dim3(W), dim3(128), sm, s, a)
  switch (NB) {
    case 8: AFQ_RATIO(8); break;
    case 12: AFQ_RATIO(12); break;
    case 16: AFQ_RATIO(16); break;
    case 20: AFQ_RATIO(20); break;
    case 24: AFQ_RATIO(24); break;
    case 28: AFQ_RATIO(28); break;
    case 32: AFQ_RATIO(32); break;
    case 36: AFQ_RATIO(36); break;
    default: AFQ_RATIO(40); break;
  }
#undef AFQ_RATIO
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

}  // namespace afq
