// Multi-determinant trials whose determinants are occupation strings,
// trial_s[I] = eye(M)[:, occ_s[I]] (the driver's construction,
// OmegaQMC/afqmc_gto.py:1163-1174).
//
//   greens_th4.py:75-148   Ghalf_I, o_I per det; dets with
//                           |o_Ia o_Ib| < max(1e-14 max_J |o_J|, 1e-300) are
//                           masked (their Ghalf reads as zero);
//                           w_I = c_I^* o_Ia o_Ib, overlap = sum_I w_I (det
//                           chunks summed in the reference's scan order);
//                           G_s = sum_I w_I T_I^* Ghalf_I / overlap.
//   greens_th4.py:186-228  overlap only.
//   energy_th4.py:26-119   E_1b from G, E_2b = sum_I w_I E_2b,I / overlap with
//                           the rows occ_I of L as the half-rotated vectors.
//   afqmc_gto.py:728-839   the step: force bias from G, then the single-det
//                           propagation kernels.
//
// With occupation-string dets S_I = phi^T T_I^* is the row gather
// S_I[i][j] = phi[occ_I[j]][i] and T_I^* Ghalf_I scatters the rows of Ghalf_I
// to rows occ_I of G, so no dense trial product is formed.  Each (walker,
// det) pair of spins is one workgroup of two wavefronts running the register
// Gauss-Jordan kernel of the single-det Green's function.
#include <algorithm>
#include <vector>

#include "afq_bf3.h"
#include "afq_gemm.h"
#include "afq_gj.h"
#include "afq_kernels.h"
#include "afq_mfma.h"
#include "afq_multidet.h"
#include "afq_step.h"

namespace afq {

void phase_begin(afq_system* s, int ph);
void phase_end(afq_system* s, int ph);
int lowp_step_buffers(afq_system* sys, int W, bf16_t** X, bf16_t** vhs, float** t0);

namespace {

using gj::GJ_LD;
using gj::GJ_NMAX;

constexpr double MASK_RTOL = 1e-14;   // greens_th4.py:24
constexpr double MASK_ATOL = 1e-300;  // greens_th4.py:25

// LDS of one wavefront of k_md_gj: the pivot order (64 ints) and, in mode
// 1, S^-1 (N x GJ_LD complex).
size_t md_wave_lds(int mode, int NB, int N) {
  (void)NB;
  return 256 + (mode == 1 ? 16UL * N * GJ_LD : 0);
}

// Grid W * nd (det fastest), two wavefronts: wave s = spin s.  mode 0: the
// per-det overlaps; 1: also Ghalf_I = S_I^-1 phi^T; 2: also S_I^-1 itself.
template <int NB, bool DET>
__global__ __launch_bounds__(128) void k_md_gj(int mode, int W, int M, int Na, int Nb, int nd,
                                               int wlds, const cplx* __restrict__ phia,
                                               const cplx* __restrict__ phib,
                                               const int* __restrict__ occa,
                                               const int* __restrict__ occb, cplx* __restrict__ oda,
                                               cplx* __restrict__ odb, cplx* __restrict__ gha,
                                               cplx* __restrict__ ghb, cplx* __restrict__ inva,
                                               cplx* __restrict__ invb) {
  extern __shared__ __align__(16) char smem[];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int I = blockIdx.x % nd;
  const long w = blockIdx.x / nd;
  const int N = wv ? Nb : Na;
  cplx* od = (wv ? odb : oda) + w * nd + I;
  if (N == 0) {
    if (l == 0) *od = mkc(1.0, 0.0);
    return;  // no workgroup barrier in this kernel
  }
  char* base = smem + (long)wv * wlds;
  int* order = (int*)base;
  cplx* Inv = (cplx*)(order + 64);
  const cplx* phi = wv ? phib + w * M * Nb : phia + w * M * Na;
  const int* occ = (wv ? occb : occa) + (long)I * N;
  cplx a[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j)
    a[j] = (l < N && j < N) ? phi[(long)occ[j] * N + l] : mkc(0.0, 0.0);
  int mystep = 0;
  gj::wave_gj<NB, DET>(a, N, order, mystep, od);
  if (DET) return;
  if (mode == 1) {
    gj::gj_store_inverse<NB>(a, N, mystep, order, Inv);
    cplx* G = (wv ? ghb : gha) + ((long)I * W + w) * N * M;
    gj::ghalf_mfma_wave<(NB + 3) / 4>(Inv, GJ_LD, N, M, phi, G, nullptr, 0, 0, 0, nullptr);
  } else if (mode == 2) {
    cplx* iv = (wv ? invb : inva) + ((long)I * W + w) * N * N;
    gj::gj_store_inverse<NB>(a, N, mystep, order, iv, N);
  }
}

// The dets occupying orbital p, in det order: io[p nd + k] = I << 6 | i
// (occ_I[i] = p) for k < io[M nd + p].  One thread per orbital.
__global__ void k_md_occlist(int nd, int N, int M, const int* __restrict__ occ,
                             int* __restrict__ io) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  int k = 0;
  for (int I = 0; I < nd; ++I)
    for (int i = 0; i < N; ++i)
      if (occ[(long)I * N + i] == p) io[(long)p * nd + k++] = I << 6 | i;
  io[(long)M * nd + p] = k;
}

// B_s[w][p][j] = sum_I z_I S_I^-1[i][j] (p = occ_I[i], kept dets only) /
// overlap, the dets summed per chunk and the chunks in order (a chunk with
// no det on p adds nothing: the sums are those of the loop over all dets).  Then
// G_s = T^* Ghalf summed over dets = B_s phi_s^T.  One thread per element of
// both spins' B, grid-stride over W M (Na + Nb).
__global__ __launch_bounds__(256) void k_md_bmat(long n, int W, int M, int Na, int Nb, int nd,
                                                 int chunk, const int* __restrict__ ioa,
                                                 const int* __restrict__ iob,
                                                 const cplx* __restrict__ z,
                                                 const int* __restrict__ keep,
                                                 const cplx* __restrict__ ovlp,
                                                 const cplx* __restrict__ inva,
                                                 const cplx* __restrict__ invb,
                                                 cplx* __restrict__ Ba, cplx* __restrict__ Bb) {
  const long per = (long)M * (Na + Nb);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const long w = e / per;
    long r = e % per;
    const bool sb = r >= (long)M * Na;
    if (sb) r -= (long)M * Na;
    const int N = sb ? Nb : Na;
    const int p = (int)(r / N), j = (int)(r % N);
    const int* io = sb ? iob : ioa;
    const cplx* inv = sb ? invb : inva;
    const int* el = io + (long)p * nd;
    const int ne = io[(long)M * nd + p];
    cplx acc = mkc(0.0, 0.0), cs = mkc(0.0, 0.0);
    int cur = -1;
    for (int k = 0; k < ne; ++k) {
      const int I = el[k] >> 6, i = el[k] & 63;
      const int ch = I / chunk;
      if (ch != cur) {
        if (cur >= 0) acc = cadd(acc, cs);
        cs = mkc(0.0, 0.0);
        cur = ch;
      }
      if (!keep[w * nd + I]) continue;
      cfma(cs, z[w * nd + I], inv[(((long)I * W + w) * N + i) * N + j]);
    }
    if (cur >= 0) acc = cadd(acc, cs);
    (sb ? Bb : Ba)[(w * M + p) * N + j] = cdiv(acc, ovlp[w]);
  }
}

// One thread per walker: masks, w_I = conj(c_I) o_Ia o_Ib and the overlap,
// summed per det chunk then across chunks as the reference's scan.
__global__ void k_md_weights(int W, int nd, int chunk, const cplx* __restrict__ oda,
                             const cplx* __restrict__ odb, const cplx* __restrict__ ci,
                             cplx* __restrict__ z, int* __restrict__ keep,
                             cplx* __restrict__ ovlp) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  double omax = 0.0;
  for (int I = 0; I < nd; ++I) {
    const cplx o = cmul(oda[(long)w * nd + I], odb[(long)w * nd + I]);
    omax = jmax(omax, hypot(o.x, o.y));
  }
  const double tol = jmax(MASK_RTOL * omax, MASK_ATOL);
  cplx acc = mkc(0.0, 0.0);
  for (int c0 = 0; c0 < nd; c0 += chunk) {
    cplx cs = mkc(0.0, 0.0);
    const int c1 = min(nd, c0 + chunk);
    for (int I = c0; I < c1; ++I) {
      const cplx oa = oda[(long)w * nd + I], ob = odb[(long)w * nd + I];
      const cplx o = cmul(oa, ob);
      const cplx wi = cmul(cmul(cconj(ci[I]), oa), ob);
      cs = cadd(cs, wi);
      const bool k = hypot(o.x, o.y) >= tol;
      keep[(long)w * nd + I] = k ? 1 : 0;
      z[(long)w * nd + I] = k ? wi : mkc(0.0, 0.0);
    }
    acc = cadd(acc, cs);
  }
  ovlp[w] = acc;
}

// G_s[w][p][q] = sum_I z_I Ghalf_I[i][q] (p = occ_I[i]) / overlap; the
// Ghalf_I of a masked det is zeroed in place (it may hold inf/NaN).
// Grid (W, 2), one thread per column q.
__global__ __launch_bounds__(256) void k_md_gfull(int W, int M, int Na, int Nb, int nd,
                                                  const int* __restrict__ occa,
                                                  const int* __restrict__ occb,
                                                  const cplx* __restrict__ z,
                                                  const int* __restrict__ keep,
                                                  const cplx* __restrict__ ovlp,
                                                  cplx* __restrict__ gha, cplx* __restrict__ ghb,
                                                  cplx* __restrict__ Ga, cplx* __restrict__ Gb,
                                                  int want_g) {
  const long w = blockIdx.x;
  const int sp = blockIdx.y;
  const int N = sp ? Nb : Na;
  const int* occ = sp ? occb : occa;
  cplx* gh = sp ? ghb : gha;
  cplx* G = (sp ? Gb : Ga) + w * M * M;
  const long MM = (long)M * M;
  (void)MM;
  const cplx ov = ovlp[w];
  for (int q = threadIdx.x; q < M; q += blockDim.x) {
    if (want_g)
      for (int p = 0; p < M; ++p) G[(long)p * M + q] = mkc(0.0, 0.0);
    for (int I = 0; I < nd; ++I) {
      cplx* g = gh + ((long)I * W + w) * N * M;
      if (!keep[w * nd + I]) {
        for (int i = 0; i < N; ++i) g[(long)i * M + q] = mkc(0.0, 0.0);
        continue;
      }
      if (!want_g) continue;
      const cplx zi = z[w * nd + I];
      for (int i = 0; i < N; ++i) {
        const int p = occ[(long)I * N + i];
        cfma(G[(long)p * M + q], zi, g[(long)i * M + q]);
      }
    }
    if (want_g)
      for (int p = 0; p < M; ++p) G[(long)p * M + q] = cdiv(G[(long)p * M + q], ov);
  }
}

// Gt[w][q][p] = Ga[w][p][q] + Gb[w][p][q]  (the force bias contracts L_pq G_qp;
// Gb NULL: Ga already holds the sum)
__global__ void k_md_gt(long n, int M, const cplx* __restrict__ Ga, const cplx* __restrict__ Gb,
                        cplx* __restrict__ Gt) {
  const long MM = (long)M * M;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const long w = e / MM, r = e % MM;
    const int q = (int)(r / M), p = (int)(r % M);
    const long src = w * MM + (long)p * M + q;
    Gt[e] = Gb ? cadd(Ga[src], Gb[src]) : Ga[src];
  }
}

// The packed triangle straight from B_s and phi_s (no M x M G): with
// G = sum_s B_s phi_s^T, S = G + G^T = sum_s [B_s | phi_s] [phi_s | B_s]^T,
// formed over the upper-triangle 16 x 16 blocks on FP64 MFMA; packed column
// c(p, q) (row-major upper triangle, as colmap) gets S[p][q] for p < q and
// S[p][p] / 2 = G[p][p].  One wavefront per (walker, block pair), four per
// workgroup; consecutive k-steps alternate between two accumulator sets.
// Output: FP64 Gpk[w][c] (gpk) or split-BF16 planes [2][2W][Kp], row
// (w, re/im), zero for P <= c < Kp (gbf).
__global__ __launch_bounds__(256) void k_md_gsym(int W, int M, int Na, int Nb, int RT, int P,
                                                 int Kp, const cplx* __restrict__ Ba,
                                                 const cplx* __restrict__ Bb,
                                                 const cplx* __restrict__ phia,
                                                 const cplx* __restrict__ phib,
                                                 cplx* __restrict__ gpk, bf16_t* __restrict__ gbf) {
  const int npair = RT * (RT + 1) / 2;
  const long gw = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (gw >= (long)W * npair) return;  // whole wavefront; no barrier in this kernel
  const long w = gw / npair;
  const int pr = (int)(gw % npair);
  const int l = threadIdx.x & 63, li = l & 15, lq = l >> 4;
  const long plane = 2L * W * Kp;
  if (gbf && pr == 0 && l < Kp - P) {
    const long rr = 2 * w * Kp + P + l;
    gbf[rr] = gbf[rr + Kp] = gbf[plane + rr] = gbf[plane + rr + Kp] = 0;
  }
  int PB = 0, rem = pr;  // pair -> block row PB <= block column QB
  while (rem >= RT - PB) {
    rem -= RT - PB;
    ++PB;
  }
  const int QB = PB + rem;
  const int pa = 16 * PB + li, qa = 16 * QB + li;  // A row (lane i), B column (lane j)
  gj::d4v re[2], im[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) re[t] = im[t] = gj::d4v{0.0, 0.0, 0.0, 0.0};
  for (int sp = 0; sp < 2; ++sp) {
    const int N = sp ? Nb : Na;
    if (N == 0) continue;
    const cplx* Bs = (sp ? Bb : Ba) + w * M * N;
    const cplx* ph = (sp ? phib : phia) + w * M * N;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const cplx* X = (h ? ph : Bs) + (long)min(pa, M - 1) * N;  // row of block PB
      const cplx* Y = (h ? Bs : ph) + (long)min(qa, M - 1) * N;  // row of block QB
      for (int k0 = 0; k0 < N; k0 += 8) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int j = k0 + 4 * t + lq;
          const bool ok = j < N;
          const cplx a = ok && pa < M ? X[j] : mkc(0.0, 0.0);
          const cplx b = ok && qa < M ? Y[j] : mkc(0.0, 0.0);
          re[t] = gj::mfma64(a.x, b.x, re[t]);
          re[t] = gj::mfma64(-a.y, b.y, re[t]);
          im[t] = gj::mfma64(a.x, b.y, im[t]);
          im[t] = gj::mfma64(a.y, b.x, im[t]);
        }
      }
    }
  }
  // D[i][j] in lane j + 16 (i % 4), register i / 4
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int p = 16 * PB + 4 * r + lq, q = 16 * QB + li;
    if (p >= M || q >= M || p > q) continue;
    const double f = p == q ? 0.5 : 1.0;
    const cplx v = mkc(f * (re[0][r] + re[1][r]), f * (im[0][r] + im[1][r]));
    const long c = (long)p * M - (long)p * (p - 1) / 2 + (q - p);
    if (gbf) {
      const long rr = 2 * w * Kp + c, ri = rr + Kp;
      bf16_t hh, lo;
      split2(v.x, hh, lo);
      gbf[rr] = hh;
      gbf[plane + rr] = lo;
      split2(v.y, hh, lo);
      gbf[ri] = hh;
      gbf[plane + ri] = lo;
    } else {
      gpk[w * P + c] = v;
    }
  }
}

// rchol_I[g][i][q] = L[g][occ_I[i]][q]
__global__ void k_md_rchol(long n, int M, int N, const int* __restrict__ occ,
                           const double* __restrict__ L, double* __restrict__ out) {
  const long NM = (long)N * M;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const long g = e / NM, r = e % NM;
    const int i = (int)(r / M), q = (int)(r % M);
    out[e] = L[(g * M + occ[i]) * M + q];
  }
}

// E2[I][w] = e_coul - e_exch of det I
__global__ void k_md_e2b_store(int W, const cplx* __restrict__ ec, const cplx* __restrict__ ex,
                               cplx* __restrict__ e2) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w < W) e2[w] = csub(ec[w], ex[w]);
}

// e2b = (sum over chunks of sum_I z_I E2_I) / overlap; e_tot = e1b + e2b
__global__ void k_md_e2b_final(int W, int nd, int chunk, const cplx* __restrict__ z,
                               const cplx* __restrict__ e2, const cplx* __restrict__ ovlp,
                               const cplx* __restrict__ e1b, cplx* __restrict__ etot,
                               cplx* __restrict__ e2b) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  cplx acc = mkc(0.0, 0.0);
  for (int c0 = 0; c0 < nd; c0 += chunk) {
    cplx cs = mkc(0.0, 0.0);
    for (int I = c0; I < min(nd, c0 + chunk); ++I)
      cs = cadd(cs, cmul(z[(long)w * nd + I], e2[(long)I * W + w]));
    acc = cadd(acc, cs);
  }
  const cplx v = cdiv(acc, ovlp[w]);
  if (e2b) e2b[w] = v;
  etot[w] = cadd(e1b[w], v);
}

size_t a256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

size_t md_bytes(int W, int M, int Na, int Nb, int nd, int kind, bool gfull) {
  size_t o = 0;
  if (kind == MD_GHALF || kind == MD_GHALF_RAW) {
    o += a256(16UL * nd * W * Na * M);
    o += a256(16UL * nd * W * Nb * M);
  }
  if (kind == MD_INVERSE) {
    o += a256(16UL * nd * W * Na * Na) + a256(16UL * nd * W * Nb * Nb);
    o += a256(16UL * W * M * Na) + a256(16UL * W * M * Nb);
    o += 2 * a256(4UL * (nd + 1) * M);
  }
  o += 2 * a256(16UL * W * nd);        // oda, odb
  o += a256(16UL * W * nd);            // z
  o += a256(4UL * W * nd);             // keep
  o += a256(16UL * W);                 // ovlp
  if (gfull) o += 2 * a256(16UL * W * M * M);
  return o;
}

MdBufs md_carve(void* base, int W, int M, int Na, int Nb, int nd, int kind, bool gfull) {
  MdBufs b;
  char* p = (char*)base;
  auto take = [&](size_t n) {
    char* r = p;
    p += a256(n);
    return r;
  };
  if (kind == MD_GHALF || kind == MD_GHALF_RAW) {
    b.gha = (double*)take(16UL * nd * W * Na * M);
    b.ghb = (double*)take(16UL * nd * W * Nb * M);
  }
  if (kind == MD_INVERSE) {
    b.inva = (double*)take(16UL * nd * W * Na * Na);
    b.invb = (double*)take(16UL * nd * W * Nb * Nb);
    b.Ba = (double*)take(16UL * W * M * Na);
    b.Bb = (double*)take(16UL * W * M * Nb);
    b.ioa = (int*)take(4UL * (nd + 1) * M);
    b.iob = (int*)take(4UL * (nd + 1) * M);
  }
  b.oda = (double*)take(16UL * W * nd);
  b.odb = (double*)take(16UL * W * nd);
  b.z = (double*)take(16UL * W * nd);
  b.keep = (int*)take(4UL * W * nd);
  b.ovlp = (double*)take(16UL * W);
  if (gfull) {
    b.Ga = (double*)take(16UL * W * M * M);
    b.Gb = (double*)take(16UL * W * M * M);
  }
  return b;
}

int md_check_occ(hipStream_t s, int nd, int N, int M, const int32_t* occ) {
  if (nd == 0 || N == 0) return AFQ_OK;
  if (!occ) return fail(AFQ_EINVAL, "occupation strings are NULL");
  std::vector<int32_t> h((size_t)nd * N);
  AFQ_CHECK_HIP(hipMemcpyAsync(h.data(), occ, h.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  AFQ_CHECK_HIP(hipStreamSynchronize(s));
  for (int I = 0; I < nd; ++I) {
    std::vector<int32_t> r(h.begin() + (size_t)I * N, h.begin() + (size_t)(I + 1) * N);
    for (int32_t v : r)
      if (v < 0 || v >= M) return fail(AFQ_EINVAL, "det %d: orbital index %d outside [0, %d)", I, v, M);
    std::sort(r.begin(), r.end());
    if (std::adjacent_find(r.begin(), r.end()) != r.end())
      return fail(AFQ_EINVAL, "det %d repeats an orbital", I);
  }
  return AFQ_OK;
}

int md_greens(hipStream_t s, int W, int M, int Na, int Nb, int nd, int chunk, const double* phia,
              const double* phib, const int* occa, const int* occb, const double* ci, int kind,
              bool gsum, const MdBufs& b) {
  if (W <= 0) return AFQ_OK;
  const int Nmax = std::max(Na, Nb);
  if (Nmax > GJ_NMAX) return fail(AFQ_EUNSUPPORTED, "multi-det: nocc %d > %d", Nmax, GJ_NMAX);
  if (gsum && kind == MD_GHALF) return fail(AFQ_EINVAL, "multi-det: summed G needs MD_INVERSE");
  const int mode = (kind == MD_GHALF || kind == MD_GHALF_RAW) ? 1 : kind == MD_INVERSE ? 2 : 0;
  // register row length: N rounded up to a multiple of 4 (every padding
  // column is a full update per elimination step)
  const int NB = Nmax <= 8 ? 8 : (Nmax + 3) / 4 * 4;
  const size_t wl = md_wave_lds(mode, NB, Nmax);
  const size_t sm = 2 * wl;
  if (sm > max_lds_bytes()) return fail(AFQ_EUNSUPPORTED, "multi-det: LDS %zu", sm);
  if ((long)W * nd > 0x7fffffffL) return fail(AFQ_EUNSUPPORTED, "multi-det: grid too large");
  const dim3 grid((unsigned)((long)W * nd));
#define AFQ_MDGJ(NBV)                                                                          \
  do {                                                                                            \
    if (mode == 0)                                                                                \
      hipLaunchKernelGGL((k_md_gj<NBV, true>), grid, dim3(128), sm, s, mode, W, M, Na, Nb, nd,    \
                         (int)wl, (const cplx*)phia, (const cplx*)phib, occa, occb, (cplx*)b.oda, \
                         (cplx*)b.odb, (cplx*)b.gha, (cplx*)b.ghb, (cplx*)b.inva, (cplx*)b.invb); \
    else                                                                                          \
      hipLaunchKernelGGL((k_md_gj<NBV, false>), grid, dim3(128), sm, s, mode, W, M, Na, Nb, nd,   \
                         (int)wl, (const cplx*)phia, (const cplx*)phib, occa, occb, (cplx*)b.oda, \
                         (cplx*)b.odb, (cplx*)b.gha, (cplx*)b.ghb, (cplx*)b.inva, (cplx*)b.invb); \
  } while (0)
  switch (NB) {
    case 8: AFQ_MDGJ(8); break;
    case 12: AFQ_MDGJ(12); break;
    case 16: AFQ_MDGJ(16); break;
    case 20: AFQ_MDGJ(20); break;
    case 24: AFQ_MDGJ(24); break;
    case 28: AFQ_MDGJ(28); break;
    case 32: AFQ_MDGJ(32); break;
    case 36: AFQ_MDGJ(36); break;
    default: AFQ_MDGJ(40); break;
  }
#undef AFQ_MDGJ
  AFQ_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_md_weights, dim3(cdiv_int(W, 128)), dim3(128), 0, s, W, nd,
                     std::max(1, chunk), (const cplx*)b.oda, (const cplx*)b.odb, (const cplx*)ci,
                     (cplx*)b.z, b.keep, (cplx*)b.ovlp);
  AFQ_CHECK_LAUNCH();
  if (kind == MD_GHALF) {
    hipLaunchKernelGGL(k_md_gfull, dim3(W, 2), dim3(256), 0, s, W, M, Na, Nb, nd, occa, occb,
                       (const cplx*)b.z, b.keep, (const cplx*)b.ovlp, (cplx*)b.gha, (cplx*)b.ghb,
                       (cplx*)b.Ga, (cplx*)b.Gb, b.Ga ? 1 : 0);
    AFQ_CHECK_LAUNCH();
  } else if (kind == MD_INVERSE) {
    const int nio = cdiv_int(M, 256);
    hipLaunchKernelGGL(k_md_occlist, dim3(nio), dim3(256), 0, s, nd, Na, M, occa, b.ioa);
    AFQ_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_md_occlist, dim3(nio), dim3(256), 0, s, nd, Nb, M, occb, b.iob);
    AFQ_CHECK_LAUNCH();
    const long n = (long)W * M * (Na + Nb);
    if (n > 0) {
      hipLaunchKernelGGL(k_md_bmat, dim3((unsigned)std::min<long>(cdiv_int(n, 256), 16384)),
                         dim3(256), 0, s, n, W, M, Na, Nb, nd, std::max(1, chunk), b.ioa, b.iob,
                         (const cplx*)b.z, b.keep, (const cplx*)b.ovlp, (const cplx*)b.inva,
                         (const cplx*)b.invb, (cplx*)b.Ba, (cplx*)b.Bb);
      AFQ_CHECK_LAUNCH();
    }
    if (b.Ga) return md_g_from_b(s, W, M, Na, Nb, phia, phib, gsum, b);
  }
  return AFQ_OK;
}

int md_g_from_b(hipStream_t s, int W, int M, int Na, int Nb, const double* phia,
                const double* phib, bool gsum, const MdBufs& b) {
  // G_s[w][p][q] = sum_j B_s[w][p][j] phi_s[w][q][j]; an empty spin gives 0
  for (int sp = 0; sp < 2; ++sp) {
    const int N = sp ? Nb : Na;
    GemmArgs g;
    g.m = M; g.n = M; g.k = N; g.batch = W;
    g.A = sp ? b.Bb : b.Ba; g.lda_m = N; g.lda_k = 1; g.sA = (long)M * N; g.a_cplx = true;
    g.B = sp ? phib : phia; g.ldb_k = 1; g.ldb_n = N; g.sB = (long)M * N; g.b_cplx = true;
    g.C = (cplx*)((sp && !gsum) ? b.Gb : b.Ga); g.ldc = M; g.sC = (long)M * M;
    g.beta = (sp && gsum) ? 1.0 : 0.0;
    int rc = gemm_fp64(s, g);
    if (rc) return rc;
  }
  return AFQ_OK;
}

// FP64 multi-det local energy given the buffers of md_greens(MD_GHALF) with G.
// scratch: E2 (nd W complex) | rchol_a | rchol_b | X (W G complex) | ec | ex
//          | exchange slab (gc N W N complex)
// lp: the split-BF16 operands of a BF16x3 system (each determinant's Coulomb
// and exchange then run on the split-BF16 kernels, as in the single-det
// BF16x3 estimator; NULL: FP64 two-body terms)
static bool md_bf3_exchange(const Bf3System* lp) { return lp && bf3_exchange_supported(*lp); }

static size_t md_energy_scratch(int W, int M, int Na, int Nb, int G, int nd, int* gc_out,
                                const Bf3System* lp = nullptr) {
  const long N = std::max(Na, Nb);
  long gc = N > 0 ? std::max(1L, std::min((long)G, (256L << 20) / (16L * N * W * N))) : 1;
  if (gc_out) *gc_out = (int)gc;
  const size_t slab = md_bf3_exchange(lp) ? a256(bf3_md_two_body_bytes(*lp, W))
                                          : a256(16UL * gc * N * W * N);
  return a256(16UL * nd * W) + a256(8UL * G * Na * M) + a256(8UL * G * Nb * M) +
         a256(2 * 16UL * W * G) + 2 * a256(16UL * W) + slab;
}

static int md_energy(hipStream_t s, int W, int M, int Na, int Nb, int G, int nd, int chunk,
                     const double* h1e, const double* chol, const int* occa, const int* occb,
                     double enuc, const MdBufs& b, void* scr, double* e_tot, double* e_1b,
                     double* e_2b, const Bf3System* lp = nullptr) {
  int gc = 1;
  (void)md_energy_scratch(W, M, Na, Nb, G, nd, &gc, lp);
  const bool bx = md_bf3_exchange(lp);
  char* p = (char*)scr;
  auto take = [&](size_t n) {
    char* r = p;
    p += a256(n);
    return r;
  };
  double* E2 = (double*)take(16UL * nd * W);
  double* ra = (double*)take(8UL * G * Na * M);
  double* rb = (double*)take(8UL * G * Nb * M);
  double* X = (double*)take(2 * 16UL * W * G);  // two (W,G) slices
  double* ec = (double*)take(16UL * W);
  double* ex = (double*)take(16UL * W);
  double* slab = (double*)take(bx ? bf3_md_two_body_bytes(*lp, W)
                                  : 16UL * gc * std::max(Na, Nb) * W * std::max(Na, Nb));
  int rc = e1b_fullg_launch(s, W, M, h1e, b.Ga, b.Gb, enuc, e_1b);
  if (rc) return rc;
  for (int I = 0; I < nd; ++I) {
    const double* ga = b.gha + 2L * I * W * Na * M;
    const double* gb = b.ghb + 2L * I * W * Nb * M;
    if (bx) {
      if ((rc = bf3_md_two_body(s, *lp, W, occa + (long)I * Na, occb + (long)I * Nb, chol, ga, gb,
                                slab, ec, ex)))
        return rc;
      hipLaunchKernelGGL(k_md_e2b_store, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W,
                         (const cplx*)ec, (const cplx*)ex, (cplx*)E2 + (long)I * W);
      AFQ_CHECK_LAUNCH();
      continue;
    }
    if (Na && G) {
      const long n = (long)G * Na * M;
      hipLaunchKernelGGL(k_md_rchol, dim3(2048), dim3(256), 0, s, n, M, Na, occa + (long)I * Na,
                         chol, ra);
      AFQ_CHECK_LAUNCH();
    }
    if (Nb && G) {
      const long n = (long)G * Nb * M;
      hipLaunchKernelGGL(k_md_rchol, dim3(2048), dim3(256), 0, s, n, M, Nb, occb + (long)I * Nb,
                         chol, rb);
      AFQ_CHECK_LAUNCH();
    }
    if ((rc = coulomb_launch(s, W, M, Na, Nb, G, ra, rb, false, ga, gb, X, 2, ec))) return rc;
    if ((rc = exchange_fp64(s, W, M, Na, G, ga, ra, false, gc, slab, ex, false))) return rc;
    if ((rc = exchange_fp64(s, W, M, Nb, G, gb, rb, false, gc, slab, ex, true))) return rc;
    hipLaunchKernelGGL(k_md_e2b_store, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W,
                       (const cplx*)ec, (const cplx*)ex, (cplx*)E2 + (long)I * W);
    AFQ_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(k_md_e2b_final, dim3(cdiv_int(W, 256)), dim3(256), 0, s, W, nd,
                     std::max(1, chunk), (const cplx*)b.z, (const cplx*)E2, (const cplx*)b.ovlp,
                     (const cplx*)e_1b, (cplx*)e_tot, (cplx*)e_2b);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

// ---------------------------------------------------------------------------
// the prepared multi-det system
// ---------------------------------------------------------------------------
namespace {

struct MdStep {
  MdBufs b;
  double *oo, *vbias, *xs, *cfb, *cmf, *xi, *Gt, *vhs, *ovn;
};

size_t md_step_bytes(const afq_system* s, int W) {
  const long M = s->M, G = s->G;
  size_t n = md_bytes(W, s->M, s->Na, s->Nb, s->ndet, MD_INVERSE, true);
  n += a256(16UL * W);                   // oo
  n += a256(16UL * W * G * FB_MAX_SPLITS);  // vbias (split-K partials)
  n += a256(16UL * W * G);               // xs
  n += 2 * a256(16UL * W);               // cfb, cmf
  n += a256(8UL * W * G);                // xi
  n += a256(16UL * W * M * M);           // Gt
  n += a256(16UL * W * M * M);           // FP64 VHS
  n += md_bytes(W, s->M, s->Na, s->Nb, s->ndet, MD_OVERLAP, false);  // overlap-only pass
  return n;
}

int md_ensure(afq_system* s, int W, MdStep* st, MdBufs* ovb) {
  if (!s->md_work || W > s->md_cap) {
    if (s->md_work) {
      AFQ_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
      AFQ_CHECK_HIP(hipFree(s->md_work));
      s->md_work = nullptr;
    }
    const size_t n = md_step_bytes(s, W);
    if (hipMalloc(&s->md_work, n) != hipSuccess) {
      s->md_work = nullptr;
      s->md_cap = 0;
      return fail(AFQ_ENOMEM, "multi-det workspace of %zu bytes", n);
    }
    s->md_bytes = n;
    s->md_cap = W;
  }
  const int C = s->md_cap;
  const long M = s->M, G = s->G;
  char* p = (char*)s->md_work;
  st->b = md_carve(p, C, s->M, s->Na, s->Nb, s->ndet, MD_INVERSE, true);
  p += md_bytes(C, s->M, s->Na, s->Nb, s->ndet, MD_INVERSE, true);
  auto take = [&](size_t n) {
    char* r = p;
    p += a256(n);
    return (double*)r;
  };
  st->oo = take(16UL * C);
  st->vbias = take(16UL * C * G * FB_MAX_SPLITS);
  st->xs = take(16UL * C * G);
  st->cfb = take(16UL * C);
  st->cmf = take(16UL * C);
  st->xi = take(8UL * C * G);
  st->Gt = take(16UL * C * M * M);
  st->vhs = take(16UL * C * M * M);
  *ovb = md_carve(p, C, s->M, s->Na, s->Nb, s->ndet, MD_OVERLAP, false);
  // the buffers are laid out for C walkers; re-carve for W keeps strides W
  st->b = md_carve(s->md_work, W, s->M, s->Na, s->Nb, s->ndet, MD_INVERSE, true);
  return AFQ_OK;
}

}  // namespace

int propagate_multidet(afq_system* sys, int W, double* phia, double* phib, double* weights,
                       double* e_hybrid, double eshift, const double* xi_in, uint64_t seed,
                       int64_t step, double* ovlp_new) {
  hipStream_t s = sys->ctx->stream;
  const int M = sys->M, Na = sys->Na, Nb = sys->Nb, G = sys->G, nd = sys->ndet;
  MdStep st;
  MdBufs ovb;
  int rc = md_ensure(sys, W, &st, &ovb);
  if (rc) return rc;
  const double* xi = xi_in;
  if (!xi) {
    phase_begin(sys, PH_RNG);
    rc = aux_fields_launch(s, seed, step, W, G, st.xi, sys->step_dev);
    phase_end(sys, PH_RNG);
    if (rc) return rc;
    xi = st.xi;
  }
  // 1: multi-det G (Ga + Gb into st.b.Ga) and overlap
  phase_begin(sys, PH_GREENS);
  MdBufs gb = st.b;
  if (sys->chol_packed) gb.Ga = nullptr;  // the packed triangle comes from B_s and phi_s
  if (sys->md_method == 1) {
    rc = md_ratio_launch(sys, 1, W, phia, phib, gb.ovlp, gb.Ba, gb.Bb);
    if (!rc && gb.Ga) rc = md_g_from_b(s, W, M, Na, Nb, phia, phib, true, gb);
  } else {
    rc = md_greens(s, W, M, Na, Nb, nd, sys->det_chunk, phia, phib, sys->occ_a, sys->occ_b,
                   sys->ci, MD_INVERSE, true, gb);
  }
  phase_end(sys, PH_GREENS);
  if (rc) return rc;
  // 3: vbias_g = sum_pq L_g,pq (Ga + Gb)_qp: over the packed triangle on the
  //    FP64 MFMA force-bias kernel (split-K) when L is symmetric (the
  //    split-BF16 GEMM, FP32 partials, in a BF16x3 system), else one generic
  //    GEMM over (p, q)
  phase_begin(sys, PH_FORCEBIAS);
  int nsplit = 1;
  float* fb_part = nullptr;
  if (sys->chol_packed && sys->precision == AFQ_BF16X3) {
    const int P = sys->packed_P, Kp = (P + 31) / 32 * 32;
    if (!sys->md_lpk) {
      // published only once the split has been launched without error, so a
      // failed first step never leaves a plane buffer that later steps trust
      bf16_t* lpk = nullptr;
      AFQ_CHECK_HIP(hipMalloc(&lpk, (size_t)2 * G * Kp * sizeof(bf16_t)));
      if ((rc = bf3_split_rows(s, sys->chol_packed, G, P, Kp, lpk))) {
        (void)hipStreamSynchronize(s);
        (void)hipFree(lpk);
        return rc;
      }
      sys->md_lpk = lpk;
      sys->md_kp = Kp;
    }
    const int RT = (M + 15) / 16;
    const long nwv = (long)W * (RT * (RT + 1) / 2);
    hipLaunchKernelGGL(k_md_gsym, dim3((unsigned)cdiv_int(nwv, 4)), dim3(256), 0, s, W, M, Na,
                       Nb, RT, P, Kp,
                       (const cplx*)st.b.Ba, (const cplx*)st.b.Bb, (const cplx*)phia,
                       (const cplx*)phib, (cplx*)nullptr, (bf16_t*)st.Gt);
    AFQ_CHECK_LAUNCH();
    fb_part = (float*)st.vbias;
    if ((rc = bf3_packed_force_bias(s, W, G, Kp, (const bf16_t*)st.Gt, sys->md_lpk, fb_part,
                                    &nsplit)))
      return rc;
  } else if (sys->chol_packed) {
    const int P = sys->packed_P;
    const int RT = (M + 15) / 16;
    const long nwv = (long)W * (RT * (RT + 1) / 2);
    hipLaunchKernelGGL(k_md_gsym, dim3((unsigned)cdiv_int(nwv, 4)), dim3(256), 0, s, W, M, Na,
                       Nb, RT, P, 0,
                       (const cplx*)st.b.Ba, (const cplx*)st.b.Bb, (const cplx*)phia,
                       (const cplx*)phib, (cplx*)st.Gt, (bf16_t*)nullptr);
    AFQ_CHECK_LAUNCH();
    FbSplits sp{};
    const int nsp = 4;
    const long chunk = (((long)P + nsp - 1) / nsp + 15) / 16 * 16;
    for (long k0 = 0; k0 < P && sp.n < FB_MAX_SPLITS; k0 += chunk) {
      sp.gh[sp.n] = (const cplx*)st.Gt;
      sp.lh[sp.n] = sys->chol_packed;
      sp.K[sp.n] = P;
      sp.kbeg[sp.n] = (int)k0;
      sp.kend[sp.n] = (int)std::min<long>(P, k0 + chunk);
      ++sp.n;
    }
    if ((rc = fb_mfma_f64(s, W, G, sp, st.vbias))) return rc;
    nsplit = sp.n;
  } else {
    const long n = (long)W * M * M;
    hipLaunchKernelGGL(k_md_gt, dim3(4096), dim3(256), 0, s, n, M, (const cplx*)st.b.Ga,
                       (const cplx*)nullptr, (cplx*)st.Gt);
    AFQ_CHECK_LAUNCH();
    GemmArgs g;
    g.m = W; g.n = G; g.k = M * M;
    g.A = st.Gt; g.lda_m = (long)M * M; g.lda_k = 1; g.a_cplx = true;
    g.B = sys->chol; g.ldb_k = 1; g.ldb_n = (long)M * M; g.b_cplx = false;
    g.C = (cplx*)st.vbias; g.ldc = G;
    if ((rc = gemm_fp64(s, g))) return rc;
  }
  phase_end(sys, PH_FORCEBIAS);
  if (nsplit < 1) nsplit = 1;
  // 4-6: shift, VHS, expH1 . exp(VHS) . expH1 with the single-det kernels
  if (sys->precision == AFQ_BF16X3) {
    bf16_t* X = nullptr;
    bf16_t* vb = nullptr;
    float* t0 = nullptr;
    if ((rc = lowp_step_buffers(sys, W, &X, &vb, &t0))) return rc;
    const Bf3System& b = *(const Bf3System*)sys->chol_lp;
    phase_begin(sys, PH_SHIFT);
    rc = shift_launch(s, W, G, fb_part ? nullptr : st.vbias, sys->mf_shift, xi, sys->dt,
                      sys->fbbound, nullptr, nullptr, st.cfb, st.cmf, nsplit, fb_part, X, b.Gp);
    phase_end(sys, PH_SHIFT);
    if (rc) return rc;
    phase_begin(sys, PH_VHS);
    rc = bf3_vhs(s, b, W, X, sys->dt, vb);
    phase_end(sys, PH_VHS);
    if (rc) return rc;
    phase_begin(sys, PH_TAYLOR);
    rc = bf3_propagate(s, b, W, vb, t0, phia, phib, sys->nmax);
    phase_end(sys, PH_TAYLOR);
    if (rc) return rc;
  } else {
    phase_begin(sys, PH_SHIFT);
    rc = shift_launch(s, W, G, st.vbias, sys->mf_shift, xi, sys->dt, sys->fbbound, nullptr, st.xs,
                      st.cfb, st.cmf, nsplit);
    phase_end(sys, PH_SHIFT);
    if (rc) return rc;
    phase_begin(sys, PH_VHS);
    if (sys->chol_packed)
      rc = vhs_mfma_f64(s, W, M, G, st.xs, sys->chol_packed, sys->packed_P, sys->colmap, sys->dt,
                        st.vhs);
    else
      rc = vhs_mfma_f64(s, W, M, G, st.xs, sys->chol, M * M, nullptr, sys->dt, st.vhs);
    phase_end(sys, PH_VHS);
    if (rc) return rc;
    phase_begin(sys, PH_TAYLOR);
    if (prop_f64_fused_fits(M)) {
      rc = prop_mfma_f64(s, W, M, Na, Nb, sys->expH1, st.vhs, phia, phib, phia, phib, sys->nmax);
    } else {
      // batched-GEMM propagator; its scratch lives in the (W, M, M) G^T slot,
      // free once the force bias is formed (2 (Na + Nb + max N) <= M here)
      const long Nm = std::max(Na, Nb);
      if (2L * (Na + Nb + Nm) > M) return fail(AFQ_EUNSUPPORTED, "multi-det propagator scratch");
      double* g = st.Gt;
      double* a1 = g;
      double* a2 = a1 + 2L * W * M * Na;
      double* b1 = a2 + 2L * W * M * Na;
      double* b2 = b1 + 2L * W * M * Nb;
      double* t0 = b2 + 2L * W * M * Nb;
      double* t1 = t0 + 2L * W * M * Nm;
      rc = prop_generic_f64(s, W, M, Na, Nb, sys->expH1, st.vhs, phia, phib, a1, a2, b1, b2, t0,
                            t1, sys->nmax);
    }
    phase_end(sys, PH_TAYLOR);
    if (rc) return rc;
  }
  // 7: multi-det overlap of the propagated walkers
  phase_begin(sys, PH_OVERLAP);
  MdBufs ob = md_carve(ovb.oda, W, M, Na, Nb, nd, MD_OVERLAP, false);
  ob.ovlp = ovlp_new;
  if (sys->md_method == 1)
    rc = md_ratio_launch(sys, 0, W, phia, phib, ovlp_new, nullptr, nullptr);
  else
    rc = md_greens(s, W, M, Na, Nb, nd, sys->det_chunk, phia, phib, sys->occ_a, sys->occ_b,
                   sys->ci, MD_OVERLAP, false, ob);
  phase_end(sys, PH_OVERLAP);
  if (rc) return rc;
  // 8
  phase_begin(sys, PH_WEIGHTS);
  rc = weights_launch(s, W, weights, st.b.ovlp, ovlp_new, st.cfb, st.cmf, e_hybrid, eshift,
                      sys->dt, weights, e_hybrid, sys->eshift_dev);
  phase_end(sys, PH_WEIGHTS);
  return rc;
}

int system_energy_multidet(afq_system* sys, int W, const double* phia, const double* phib,
                           const double* h1e, double enuc, double* e_tot, double* e_1b,
                           double* e_2b) {
  hipStream_t s = sys->ctx->stream;
  const int M = sys->M, Na = sys->Na, Nb = sys->Nb, G = sys->G, nd = sys->ndet;
  const size_t nb = md_bytes(W, M, Na, Nb, nd, MD_GHALF, true);
  const Bf3System* lp =
      sys->precision == AFQ_BF16X3 ? (const Bf3System*)sys->chol_lp : nullptr;
  const size_t ne = md_energy_scratch(W, M, Na, Nb, G, nd, nullptr, lp);
  char* scr = (char*)scratch(sys->ctx, nb + ne);
  if (!scr) return AFQ_ENOMEM;
  const MdBufs b = md_carve(scr, W, M, Na, Nb, nd, MD_GHALF, true);
  int rc = md_greens(s, W, M, Na, Nb, nd, sys->det_chunk, phia, phib, sys->occ_a, sys->occ_b,
                     sys->ci, MD_GHALF, false, b);
  if (rc) return rc;
  return md_energy(s, W, M, Na, Nb, G, nd, sys->det_chunk, h1e, sys->chol, sys->occ_a,
                   sys->occ_b, enuc, b, scr + nb, e_tot, e_1b, e_2b, lp);
}

namespace {

// greens.py:353-358: non-finite Ghalf entries (the singular S of a det the
// walker is orthogonal to) become 0
__global__ void k_md_sanitize(long n, cplx* __restrict__ g) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n;
       e += (long)gridDim.x * blockDim.x) {
    const cplx x = g[e];
    if (!(isfinite(x.x) && isfinite(x.y))) g[e] = mkc(0.0, 0.0);
  }
}

// greens.py:360-364: overlap = sum_I conj(c_I) o_Ia o_Ib over all dets in
// order (no mask); per-det overlaps out in the (ndet, W) layout
__global__ void k_md_fb_overlaps(int W, int nd, const cplx* __restrict__ ci,
                                 const cplx* __restrict__ oda, const cplx* __restrict__ odb,
                                 cplx* __restrict__ ovlp, cplx* __restrict__ oa_all,
                                 cplx* __restrict__ ob_all) {
  const int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= W) return;
  cplx acc = mkc(0.0, 0.0);
  for (int I = 0; I < nd; ++I) {
    const cplx a = oda[(long)w * nd + I], b = odb[(long)w * nd + I];
    acc = cadd(acc, cmul(cmul(cconj(ci[I]), a), b));
    if (oa_all) oa_all[(long)I * W + w] = a;
    if (ob_all) ob_all[(long)I * W + w] = b;
  }
  ovlp[w] = acc;
}

}  // namespace

}  // namespace afq

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int afq_greens_function_multidet(afq_ctx* ctx, int W, int M, int Na, int Nb, int nd,
                                 const double* phia, const double* phib, const int32_t* occ_up,
                                 const int32_t* occ_dn, const double* ci, int det_chunk_size,
                                 double* Ga, double* Gb, double* overlap) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && Na <= M && Nb <= M && nd > 0, "bad shape");
  AFQ_REQUIRE((phia || Na == 0) && (phib || Nb == 0) && ci && Ga && Gb && overlap,
              "NULL argument");
  hipStream_t s = ctx->stream;
  int rc = afq::md_check_occ(s, nd, Na, M, occ_up);
  if (!rc) rc = afq::md_check_occ(s, nd, Nb, M, occ_dn);
  if (rc || W == 0) return rc;
  const size_t n = afq::md_bytes(W, M, Na, Nb, nd, afq::MD_INVERSE, false);
  void* scr = afq::scratch(ctx, n);
  if (!scr) return AFQ_ENOMEM;
  afq::MdBufs b = afq::md_carve(scr, W, M, Na, Nb, nd, afq::MD_INVERSE, false);
  b.Ga = Ga;
  b.Gb = Gb;
  b.ovlp = overlap;
  return afq::md_greens(s, W, M, Na, Nb, nd, det_chunk_size, phia, phib, occ_up, occ_dn, ci,
                        afq::MD_INVERSE, false, b);
}

int afq_greens_function_multidet_overlap_only(afq_ctx* ctx, int W, int M, int Na, int Nb, int nd,
                                              const double* phia, const double* phib,
                                              const int32_t* occ_up, const int32_t* occ_dn,
                                              const double* ci, int det_chunk_size,
                                              double* overlap) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && Na <= M && Nb <= M && nd > 0, "bad shape");
  AFQ_REQUIRE((phia || Na == 0) && (phib || Nb == 0) && ci && overlap, "NULL argument");
  hipStream_t s = ctx->stream;
  int rc = afq::md_check_occ(s, nd, Na, M, occ_up);
  if (!rc) rc = afq::md_check_occ(s, nd, Nb, M, occ_dn);
  if (rc || W == 0) return rc;
  void* scr = afq::scratch(ctx, afq::md_bytes(W, M, Na, Nb, nd, afq::MD_OVERLAP, false));
  if (!scr) return AFQ_ENOMEM;
  afq::MdBufs b = afq::md_carve(scr, W, M, Na, Nb, nd, afq::MD_OVERLAP, false);
  b.ovlp = overlap;
  return afq::md_greens(s, W, M, Na, Nb, nd, det_chunk_size, phia, phib, occ_up, occ_dn, ci,
                        afq::MD_OVERLAP, false, b);
}

int afq_greens_function_multidet_force_bias(afq_ctx* ctx, int W, int M, int Na, int Nb, int nd,
                                            const double* phia, const double* phib,
                                            const int32_t* occ_up, const int32_t* occ_dn,
                                            const double* ci, double* Ghalfa_all,
                                            double* Ghalfb_all, double* overlap,
                                            double* ovlp_a_all, double* ovlp_b_all) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && Na <= M && Nb <= M && nd > 0, "bad shape");
  AFQ_REQUIRE((phia || Na == 0) && (phib || Nb == 0) && ci && overlap, "NULL argument");
  AFQ_REQUIRE((Ghalfa_all || Na == 0) && (Ghalfb_all || Nb == 0), "NULL Ghalf output");
  hipStream_t s = ctx->stream;
  int rc = afq::md_check_occ(s, nd, Na, M, occ_up);
  if (!rc) rc = afq::md_check_occ(s, nd, Nb, M, occ_dn);
  if (rc || W == 0) return rc;
  // the Ghalf planes are the caller's outputs; the rest comes from scratch
  const size_t n = afq::md_bytes(W, M, Na, Nb, nd, afq::MD_OVERLAP, false);
  void* scr = afq::scratch(ctx, n);
  if (!scr) return AFQ_ENOMEM;
  afq::MdBufs b = afq::md_carve(scr, W, M, Na, Nb, nd, afq::MD_OVERLAP, false);
  b.gha = Ghalfa_all;
  b.ghb = Ghalfb_all;
  if ((rc = afq::md_greens(s, W, M, Na, Nb, nd, 1, phia, phib, occ_up, occ_dn, ci,
                           afq::MD_GHALF_RAW, false, b)))
    return rc;
  for (int sp = 0; sp < 2; ++sp) {
    const long ne = (long)nd * W * (sp ? Nb : Na) * M;
    if (ne == 0) continue;
    hipLaunchKernelGGL(afq::k_md_sanitize, dim3((unsigned)std::min<long>((ne + 255) / 256, 8192)),
                       dim3(256), 0, s, ne, (cplx*)(sp ? Ghalfb_all : Ghalfa_all));
    AFQ_CHECK_LAUNCH();
  }
  hipLaunchKernelGGL(afq::k_md_fb_overlaps, dim3(cdiv_int(W, 128)), dim3(128), 0, s, W, nd,
                     (const cplx*)ci, (const cplx*)b.oda, (const cplx*)b.odb, (cplx*)overlap,
                     (cplx*)ovlp_a_all, (cplx*)ovlp_b_all);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int afq_local_energy_multidet(afq_ctx* ctx, int W, int M, int Na, int Nb, int G, int nd,
                              const double* h1e, const double* chol, const double* phia,
                              const double* phib, const int32_t* occ_up, const int32_t* occ_dn,
                              const double* ci, double enuc, int det_chunk_size, double* e_tot,
                              double* e_1b, double* e_2b) {
  AFQ_REQUIRE(ctx, "ctx is NULL");
  AFQ_REQUIRE(W >= 0 && M > 0 && Na >= 0 && Nb >= 0 && Na <= M && Nb <= M && nd > 0 && G >= 0,
              "bad shape");
  AFQ_REQUIRE(h1e && (chol || G == 0) && (phia || Na == 0) && (phib || Nb == 0) && ci && e_tot &&
                  e_1b,
              "NULL argument");
  hipStream_t s = ctx->stream;
  int rc = afq::md_check_occ(s, nd, Na, M, occ_up);
  if (!rc) rc = afq::md_check_occ(s, nd, Nb, M, occ_dn);
  if (rc || W == 0) return rc;
  const size_t nb = afq::md_bytes(W, M, Na, Nb, nd, afq::MD_GHALF, true);
  const size_t ne = afq::md_energy_scratch(W, M, Na, Nb, G, nd, nullptr);
  char* scr = (char*)afq::scratch(ctx, nb + ne);
  if (!scr) return AFQ_ENOMEM;
  afq::MdBufs b = afq::md_carve(scr, W, M, Na, Nb, nd, afq::MD_GHALF, true);
  if ((rc = afq::md_greens(s, W, M, Na, Nb, nd, det_chunk_size, phia, phib, occ_up, occ_dn, ci,
                           afq::MD_GHALF, false, b)))
    return rc;
  return afq::md_energy(s, W, M, Na, Nb, G, nd, det_chunk_size, h1e, chol, occ_up, occ_dn, enuc,
                        b, scr + nb, e_tot, e_1b, e_2b);
}

int afq_system_create_multidet(afq_ctx* ctx, afq_system** out, int M, int Na, int Nb, int G,
                               const double* chol, const double* expH1, const double* mf,
                               int nd, const int32_t* occ_up, const int32_t* occ_dn,
                               const double* ci, double dt, double fbbound, int nmax,
                               int det_chunk_size, int precision) {
  AFQ_REQUIRE(ctx && out, "ctx/out is NULL");
  AFQ_REQUIRE(M > 0 && Na >= 0 && Nb >= 0 && Na <= M && Nb <= M && G >= 0 && nd > 0, "bad shape");
  AFQ_REQUIRE(std::max(Na, Nb) <= afq::gj::GJ_NMAX, "multi-det: nocc > %d is not supported",
              afq::gj::GJ_NMAX);
  AFQ_REQUIRE(dt > 0 && nmax >= 0 && det_chunk_size > 0, "bad dt / nmax / det_chunk_size");
  AFQ_REQUIRE(precision == AFQ_FP64 || precision == AFQ_BF16X3,
              "multi-det: precision %d is not built (use FP64 or BF16X3)", precision);
  AFQ_REQUIRE((chol || G == 0) && expH1 && (mf || G == 0) && ci, "NULL system array");
  hipStream_t s = ctx->stream;
  int rc = afq::md_check_occ(s, nd, Na, M, occ_up);
  if (!rc) rc = afq::md_check_occ(s, nd, Nb, M, occ_dn);
  if (rc) return rc;
  afq_system* sy = new afq_system();
  sy->ctx = ctx;
  sy->M = M; sy->Na = Na; sy->Nb = Nb; sy->G = G;
  sy->chol = chol; sy->expH1 = expH1; sy->mf_shift = mf;
  sy->dt = dt; sy->fbbound = fbbound; sy->nmax = nmax; sy->precision = precision;
  sy->ndet = nd;
  sy->det_chunk = det_chunk_size;
  if (hipMalloc(&sy->occ_a, sizeof(int) * (size_t)nd * std::max(Na, 1)) != hipSuccess ||
      hipMalloc(&sy->occ_b, sizeof(int) * (size_t)nd * std::max(Nb, 1)) != hipSuccess ||
      hipMalloc(&sy->ci, 16UL * nd) != hipSuccess) {
    afq_system_destroy(sy);
    return afq::fail(AFQ_ENOMEM, "multi-det trial arrays");
  }
  hipError_t e = hipSuccess;
  if (Na) e = hipMemcpyAsync(sy->occ_a, occ_up, sizeof(int) * (size_t)nd * Na, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess && Nb)
    e = hipMemcpyAsync(sy->occ_b, occ_dn, sizeof(int) * (size_t)nd * Nb, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(sy->ci, ci, 16UL * nd, hipMemcpyDeviceToDevice, s);
  if (e != hipSuccess) {
    afq_system_destroy(sy);
    return afq::hip_fail(e, "multi-det trial copy");
  }
  rc = afq::prepare_fp64(sy);
  if (!rc) rc = afq::system_prepare_lowp(sy);
  bool ratio = false;
  if (!rc) rc = afq::md_ratio_prepare(sy, &ratio);
  if (rc) {
    afq_system_destroy(sy);
    return rc;
  }
  // the step uses ratio updates whenever every det is within four orbitals
  // of det 0 and the kernel's LDS fits
  sy->md_method = (ratio && afq::md_ratio_fits(sy)) ? 1 : 0;
  *out = sy;
  return AFQ_OK;
}

int afq_system_set_multidet_method(afq_system* sys, int method) {
  AFQ_REQUIRE(sys && sys->ndet > 0, "not a multi-determinant system");
  AFQ_REQUIRE(method == 0 || method == 1, "method %d (0 explicit, 1 ratio)", method);
  if (method == 1 && !(sys->md_ratio && afq::md_ratio_fits(sys)))
    return afq::fail(AFQ_EUNSUPPORTED,
                     "ratio updates need every det within four orbitals of det 0");
  sys->md_method = method;
  return AFQ_OK;
}

int afq_system_get_multidet_method(afq_system* sys) { return sys ? sys->md_method : -1; }

}  // extern "C"
