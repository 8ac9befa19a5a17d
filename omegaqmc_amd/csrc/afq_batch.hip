// Many independent systems stepped together: the driver step of
// OmegaQMC/afqmc_gto.py:1383-1451 (QR every stabilize_freq steps before
// propagating, the fused phaseless step, weight clip from step 2 on, comb
// every pop_control_freq steps, hybrid-energy accumulators) for each system
// of an uneven batch, with its own walker count, precision and seed.
//
// The reference runs one system per driver, one eager dispatch at a time.
// Here the systems are spread over a few HIP streams (each with its own
// context, so the QR scratch of concurrent streams never aliases), and one
// driver cycle -- lcm(stabilize_freq, pop_control_freq) steps -- of every
// system is captured into a single hipGraph whose per-stream work forms
// parallel branches.  Everything that changes from step to step is read from
// device memory by the captured kernels: the step number (auxiliary fields,
// comb offset, the step-1 clip exception) comes from a per-stream counter
// that the graph itself increments, the energy shift from a per-system
// device value the host updates between launches.  The comb's ping-pong
// walker buffers alternate from cycle to cycle, so two graphs are captured
// (starting in buffer 0 and in buffer 1) and launched alternately.  The
// graph holds no memset nodes.
#include <numeric>
#include <vector>

#include "afq_kernels.h"

namespace afq {

namespace {

__global__ void k_step_inc(int64_t* __restrict__ step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

// phi[w] = trial (M x N complex) for every walker; weights 1, e_hybrid 0,
// overlap 1 (afqmc_gto.py:76-123)
__global__ void k_init_walkers(int W, long MN, const cplx* __restrict__ trial,
                               cplx* __restrict__ phi) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < (long)W * MN;
       i += (long)gridDim.x * blockDim.x)
    phi[i] = trial[i % MN];
}

__global__ void k_init_scalars(int W, double* __restrict__ w, cplx* __restrict__ eh,
                               cplx* __restrict__ ov, double* __restrict__ acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < W) {
    w[i] = 1.0;
    eh[i] = mkc(0.0, 0.0);
    ov[i] = mkc(1.0, 0.0);
  }
  if (i < 2) acc[i] = 0.0;
}

__global__ void k_zero_acc(double* __restrict__ acc) {
  if (threadIdx.x < 2) acc[threadIdx.x] = 0.0;
}

int lcm_int(int a, int b) { return a / std::gcd(a, b) * b; }

}  // namespace

struct BatchSys {
  afq_system* sys = nullptr;
  int W = 0, stream = 0;
  uint64_t seed = 0;
  double* pa[2] = {};
  double* pb[2] = {};
  double* w[2] = {};
  double* eh[2] = {};
  double* ov = nullptr;
  double* ldr = nullptr;
  double* acc = nullptr;
  int32_t* idx = nullptr;
  double* cum = nullptr;  // comb running sums when W exceeds LDS
};

}  // namespace afq

struct afq_batch {
  afq_ctx* ctx = nullptr;                  // caller's context (its stream orders the batch)
  std::vector<afq::BatchSys> s;
  std::vector<afq_ctx*> sctx;              // one per stream, own scratch
  hipStream_t origin = nullptr;            // capture / launch stream
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::vector<hipEvent_t> ev_join;
  int64_t* step_dev = nullptr;             // one counter per stream
  double* eshift_dev = nullptr;            // per system
  double* r_dev = nullptr;                 // per system comb offset
  uint64_t* seeds_dev = nullptr;
  double* clip_tmp = nullptr;              // 32 doubles per stream
  int stab = 5, pop = 10, cycle = 10;
  int parity = 0;                          // which ping-pong buffer holds the walkers
  int64_t step = 0;                        // steps done
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  std::vector<void*> owned;
};

namespace afq {

namespace {

void* take(afq_batch* b, size_t bytes, int* rc) {
  void* p = nullptr;
  if (*rc) return nullptr;
  if (hipMalloc(&p, align256(bytes)) != hipSuccess) {
    *rc = fail(AFQ_ENOMEM, "batch buffer of %zu bytes", bytes);
    return nullptr;
  }
  b->owned.push_back(p);
  return p;
}

// Bind the systems to the batch's stream contexts and device step inputs for
// the duration of a run; the destructor undoes it.
struct Bind {
  afq_batch* b;
  std::vector<afq_ctx*> ctx;
  explicit Bind(afq_batch* bb) : b(bb) {
    for (size_t i = 0; i < b->s.size(); ++i) {
      afq_system* y = b->s[i].sys;
      ctx.push_back(y->ctx);
      y->ctx = b->sctx[b->s[i].stream];
      y->step_dev = b->step_dev + b->s[i].stream;
      y->eshift_dev = b->eshift_dev + i;
    }
  }
  ~Bind() {
    for (size_t i = 0; i < b->s.size(); ++i) {
      b->s[i].sys->ctx = ctx[i];
      b->s[i].sys->step_dev = nullptr;
      b->s[i].sys->eshift_dev = nullptr;
    }
  }
};

// Record steps [n0 + 1, n0 + nsteps] of every system, each on its stream,
// starting from ping-pong buffer p; returns the buffer that holds the
// walkers afterwards.
int record_steps(afq_batch* b, int64_t n0, int nsteps, int p, int* p_out) {
  const int ns = (int)b->sctx.size();
  for (int j = 1; j <= nsteps; ++j) {
    const int64_t n = n0 + j;
    for (int k = 0; k < ns; ++k) {
      hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, b->sctx[k]->stream, b->step_dev + k);
      AFQ_CHECK_LAUNCH();
    }
    const bool qr = n % b->stab == 0, comb = n % b->pop == 0;
    for (size_t i = 0; i < b->s.size(); ++i) {
      BatchSys& x = b->s[i];
      afq_system* y = x.sys;
      afq_ctx* c = b->sctx[x.stream];
      hipStream_t st = c->stream;
      int rc;
      if (qr) {
        void* sc = scratch(c, ortho_scratch_bytes(x.W, y->Na, y->Nb));
        if (!sc) return AFQ_ENOMEM;
        if ((rc = ortho_launch(st, x.W, y->M, y->Na, y->Nb, x.pa[p], x.pb[p], x.ldr, nullptr, sc)))
          return rc;
      }
      if ((rc = afq_propagate_walkers(y, x.W, x.pa[p], x.pb[p], x.w[p], x.eh[p], 0.0, nullptr,
                                      x.seed, n, x.ov)))
        return rc;
      if ((rc = clip_weights_launch(st, x.W, x.w[p], 0.10, b->clip_tmp + 32 * x.stream,
                                    b->step_dev + x.stream)))
        return rc;
      const int q = comb ? 1 - p : p;
      if (comb) {
        if ((rc = comb_offsets_dev_launch(st, 1, b->seeds_dev + i, b->step_dev + x.stream,
                                          b->r_dev + i)))
          return rc;
        if ((rc = comb_launch(st, x.W, y->M, y->Na, y->Nb, x.w[p], x.pa[p], x.pb[p], 0.5, x.eh[p],
                              x.w[q], x.pa[q], x.pb[q], x.eh[q], x.idx, x.cum, b->r_dev + i)))
          return rc;
      }
      if ((rc = accumulate_launch(st, x.W, x.w[q], x.eh[q], x.acc))) return rc;
    }
    if (comb) p = 1 - p;
  }
  *p_out = p;
  return AFQ_OK;
}

int capture(afq_batch* b, int p0, hipGraphExec_t* out) {
  const int ns = (int)b->sctx.size();
  AFQ_CHECK_HIP(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
  AFQ_CHECK_HIP(hipEventRecord(b->ev_in, b->origin));
  for (int k = 0; k < ns; ++k) AFQ_CHECK_HIP(hipStreamWaitEvent(b->sctx[k]->stream, b->ev_in, 0));
  int p1 = p0;
  int rc = record_steps(b, b->step, b->cycle, p0, &p1);
  for (int k = 0; k < ns; ++k) {
    (void)hipEventRecord(b->ev_join[k], b->sctx[k]->stream);
    (void)hipStreamWaitEvent(b->origin, b->ev_join[k], 0);
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->origin, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  return AFQ_OK;
}

}  // namespace

}  // namespace afq

extern "C" {

int afq_batch_create(afq_ctx* ctx, afq_batch** out, int nsys, afq_system* const* systems,
                     const int* nwalkers, const uint64_t* seeds, int stabilize_freq,
                     int pop_control_freq, int nstreams, const int* stream_of) {
  AFQ_REQUIRE(ctx && out && systems && nwalkers && seeds, "NULL argument");
  AFQ_REQUIRE(nsys > 0 && nstreams > 0 && nstreams <= 16, "bad system / stream count");
  AFQ_REQUIRE(stabilize_freq > 0 && pop_control_freq > 0, "bad frequencies");
  afq_batch* b = new afq_batch();
  b->ctx = ctx;
  b->stab = stabilize_freq;
  b->pop = pop_control_freq;
  b->cycle = afq::lcm_int(stabilize_freq, pop_control_freq);
  int rc = AFQ_OK;
  auto bad = [&](int code) {
    afq_batch_destroy(b);
    return code;
  };
  for (int k = 0; k < nstreams && !rc; ++k) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
      return bad(afq::fail(AFQ_EHIP, "stream creation"));
    afq_ctx* c = new afq_ctx();
    c->device = ctx->device;
    c->stream = st;
    c->comb_static = ctx->comb_static;
    b->sctx.push_back(c);
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess)
      return bad(afq::fail(AFQ_EHIP, "event creation"));
    b->ev_join.push_back(ev);
  }
  if (hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess)
    return bad(afq::fail(AFQ_EHIP, "stream / event creation"));
  b->step_dev = (int64_t*)afq::take(b, sizeof(int64_t) * nstreams, &rc);
  b->eshift_dev = (double*)afq::take(b, sizeof(double) * nsys, &rc);
  b->r_dev = (double*)afq::take(b, sizeof(double) * nsys, &rc);
  b->seeds_dev = (uint64_t*)afq::take(b, sizeof(uint64_t) * nsys, &rc);
  b->clip_tmp = (double*)afq::take(b, sizeof(double) * 32 * nstreams, &rc);
  if (rc) return bad(rc);
  hipStream_t st0 = ctx->stream;
  std::vector<int64_t> zero(nstreams, 0);
  std::vector<double> zd(nsys, 0.0);
  AFQ_CHECK_HIP(hipMemcpyAsync(b->step_dev, zero.data(), sizeof(int64_t) * nstreams,
                               hipMemcpyHostToDevice, st0));
  AFQ_CHECK_HIP(hipMemcpyAsync(b->eshift_dev, zd.data(), sizeof(double) * nsys,
                               hipMemcpyHostToDevice, st0));
  AFQ_CHECK_HIP(hipMemcpyAsync(b->seeds_dev, seeds, sizeof(uint64_t) * nsys, hipMemcpyHostToDevice,
                               st0));
  const size_t lds_w = (afq::max_lds_bytes() - 1024) / sizeof(double);
  for (int i = 0; i < nsys && !rc; ++i) {
    afq_system* y = systems[i];
    const int W = nwalkers[i];
    if (!y || W <= 0 || y->ctx->device != ctx->device)
      return bad(afq::fail(AFQ_EINVAL, "batch system %d: NULL, no walkers or another device", i));
    if (!y->trial_up && y->Na > 0)
      return bad(afq::fail(AFQ_EINVAL, "batch system %d: no single-determinant trial", i));
    afq::BatchSys x;
    x.sys = y;
    x.W = W;
    x.seed = seeds[i];
    x.stream = stream_of ? stream_of[i] : i % nstreams;
    if (x.stream < 0 || x.stream >= nstreams)
      return bad(afq::fail(AFQ_EINVAL, "batch system %d: stream %d of %d", i, x.stream, nstreams));
    const size_t na = 16 * (size_t)W * y->M * y->Na, nb = 16 * (size_t)W * y->M * y->Nb;
    for (int p = 0; p < 2; ++p) {
      x.pa[p] = (double*)afq::take(b, std::max<size_t>(na, 16), &rc);
      x.pb[p] = (double*)afq::take(b, std::max<size_t>(nb, 16), &rc);
      x.w[p] = (double*)afq::take(b, 8 * (size_t)W, &rc);
      x.eh[p] = (double*)afq::take(b, 16 * (size_t)W, &rc);
    }
    x.ov = (double*)afq::take(b, 16 * (size_t)W, &rc);
    x.ldr = (double*)afq::take(b, 8 * (size_t)W, &rc);
    x.acc = (double*)afq::take(b, 16, &rc);
    x.idx = (int32_t*)afq::take(b, 4 * (size_t)W, &rc);
    if ((size_t)W > lds_w) x.cum = (double*)afq::take(b, 8 * (size_t)W, &rc);
    if (rc) return bad(rc);
    const long MNa = (long)y->M * y->Na, MNb = (long)y->M * y->Nb;
    if (MNa > 0)
      hipLaunchKernelGGL(afq::k_init_walkers, dim3(1024), dim3(256), 0, st0, W, MNa,
                         (const cplx*)y->trial_up, (cplx*)x.pa[0]);
    if (MNb > 0)
      hipLaunchKernelGGL(afq::k_init_walkers, dim3(1024), dim3(256), 0, st0, W, MNb,
                         (const cplx*)y->trial_dn, (cplx*)x.pb[0]);
    hipLaunchKernelGGL(afq::k_init_scalars, dim3(cdiv_int(std::max(W, 2), 256)), dim3(256), 0, st0,
                       W, x.w[0], (cplx*)x.eh[0], (cplx*)x.ov, x.acc);
    AFQ_CHECK_LAUNCH();
    b->s.push_back(x);
  }
  AFQ_CHECK_HIP(hipStreamSynchronize(st0));
  *out = b;
  return AFQ_OK;
}

int afq_batch_set_eshift(afq_batch* b, const double* eshift) {
  AFQ_REQUIRE(b && eshift, "NULL argument");
  AFQ_CHECK_HIP(hipMemcpyAsync(b->eshift_dev, eshift, sizeof(double) * b->s.size(),
                               hipMemcpyHostToDevice, b->ctx->stream));
  AFQ_CHECK_HIP(hipStreamSynchronize(b->ctx->stream));
  return AFQ_OK;
}

int afq_batch_run(afq_batch* b, int nsteps, int mode) {
  AFQ_REQUIRE(b && nsteps >= 0, "bad argument");
  AFQ_REQUIRE(mode >= 0 && mode <= 2, "mode must be 0 (one stream), 1 (streams) or 2 (graph)");
  if (nsteps == 0) return AFQ_OK;
  afq::Bind bind(b);
  const int ns = (int)b->sctx.size();
  // order after the caller's stream, run, and order the caller's stream after
  AFQ_CHECK_HIP(hipEventRecord(b->ev_in, b->ctx->stream));
  AFQ_CHECK_HIP(hipStreamWaitEvent(b->origin, b->ev_in, 0));
  int rc = AFQ_OK;
  if (mode == 0) {
    // every system in turn on one stream: the reference's one-driver-at-a-time
    std::vector<int> keep(b->s.size());
    afq_ctx* c0 = b->sctx[0];
    for (size_t i = 0; i < b->s.size(); ++i) {
      keep[i] = b->s[i].stream;
      b->s[i].stream = 0;
      b->s[i].sys->ctx = c0;
      b->s[i].sys->step_dev = b->step_dev;
    }
    AFQ_CHECK_HIP(hipStreamWaitEvent(c0->stream, b->ev_in, 0));
    // the other counters follow stream 0's, so later runs in any mode agree
    int p = b->parity;
    std::vector<afq_ctx*> all = b->sctx;
    b->sctx.assign(1, c0);
    rc = afq::record_steps(b, b->step, nsteps, p, &p);
    b->sctx = all;
    for (size_t i = 0; i < b->s.size(); ++i) b->s[i].stream = keep[i];
    if (!rc) {
      for (int k = 1; k < ns && !rc; ++k)
        if (hipMemcpyAsync(b->step_dev + k, b->step_dev, sizeof(int64_t),
                           hipMemcpyDeviceToDevice, c0->stream) != hipSuccess)
          rc = afq::fail(AFQ_EHIP, "step counter copy");
      b->parity = p;
      b->step += nsteps;
    }
    (void)hipEventRecord(b->ev_join[0], c0->stream);
    (void)hipStreamWaitEvent(b->ctx->stream, b->ev_join[0], 0);
    return rc;
  }
  int left = nsteps;
  // eager multi-stream steps: up to the next cycle boundary, or all of them
  auto eager = [&](int n) -> int {
    for (int k = 0; k < ns; ++k) AFQ_CHECK_HIP(hipStreamWaitEvent(b->sctx[k]->stream, b->ev_in, 0));
    int p = b->parity;
    int r = afq::record_steps(b, b->step, n, p, &p);
    if (r) return r;
    b->parity = p;
    b->step += n;
    for (int k = 0; k < ns; ++k) {
      AFQ_CHECK_HIP(hipEventRecord(b->ev_join[k], b->sctx[k]->stream));
      AFQ_CHECK_HIP(hipStreamWaitEvent(b->origin, b->ev_join[k], 0));
    }
    AFQ_CHECK_HIP(hipEventRecord(b->ev_in, b->origin));
    return (int)AFQ_OK;
  };
  if (mode == 1) {
    rc = eager(left);
    left = 0;
  } else {
    const int head = (int)((b->cycle - b->step % b->cycle) % b->cycle);
    if (head > 0) {
      rc = eager(std::min(head, left));
      left -= std::min(head, left);
    }
    // the first whole cycle runs eagerly: it sizes every workspace and
    // scratch buffer the captured kernels then point into
    if (!rc && left >= b->cycle && !b->exec[0]) {
      rc = eager(b->cycle);
      left -= b->cycle;
      if (!rc) {
        AFQ_CHECK_HIP(hipStreamSynchronize(b->origin));
        const int ncomb = b->cycle / b->pop;
        for (int p0 = 0; p0 < 2 && !rc; ++p0) {
          rc = afq::capture(b, p0, &b->exec[p0]);
          if (ncomb % 2 == 0) {
            b->exec[1] = b->exec[0];
            break;
          }
        }
      }
    }
    while (!rc && left >= b->cycle) {
      AFQ_CHECK_HIP(hipGraphLaunch(b->exec[b->parity], b->origin));
      if ((b->cycle / b->pop) % 2) b->parity ^= 1;
      b->step += b->cycle;
      left -= b->cycle;
    }
    if (!rc) AFQ_CHECK_HIP(hipEventRecord(b->ev_in, b->origin));
    if (!rc && left > 0) rc = eager(left);
  }
  (void)hipEventRecord(b->ev_out, b->origin);
  (void)hipStreamWaitEvent(b->ctx->stream, b->ev_out, 0);
  return rc;
}

int afq_batch_state(afq_batch* b, int i, double** phia, double** phib, double** weights,
                    double** e_hybrid, double** overlap, double** acc, int64_t* step) {
  AFQ_REQUIRE(b && i >= 0 && i < (int)b->s.size(), "bad batch / system index");
  const afq::BatchSys& x = b->s[i];
  const int p = b->parity;
  if (phia) *phia = x.pa[p];
  if (phib) *phib = x.pb[p];
  if (weights) *weights = x.w[p];
  if (e_hybrid) *e_hybrid = x.eh[p];
  if (overlap) *overlap = x.ov;
  if (acc) *acc = x.acc;
  if (step) *step = b->step;
  return AFQ_OK;
}

int afq_batch_copy_state(afq_batch* b, int i, double* phia, double* phib, double* weights,
                         double* e_hybrid, double* overlap, double* acc) {
  AFQ_REQUIRE(b && i >= 0 && i < (int)b->s.size(), "bad batch / system index");
  const afq::BatchSys& x = b->s[i];
  const afq_system* y = x.sys;
  const int p = b->parity;
  hipStream_t st = b->ctx->stream;
  const size_t W = x.W;
  struct { double* dst; const double* src; size_t bytes; } cp[6] = {
      {phia, x.pa[p], 16 * W * y->M * y->Na}, {phib, x.pb[p], 16 * W * y->M * y->Nb},
      {weights, x.w[p], 8 * W}, {e_hybrid, x.eh[p], 16 * W}, {overlap, x.ov, 16 * W},
      {acc, x.acc, 16}};
  for (auto& c : cp)
    if (c.dst && c.bytes)
      AFQ_CHECK_HIP(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToDevice, st));
  return AFQ_OK;
}

int afq_batch_zero_accumulators(afq_batch* b) {
  AFQ_REQUIRE(b, "NULL argument");
  for (const afq::BatchSys& x : b->s)
    hipLaunchKernelGGL(afq::k_zero_acc, dim3(1), dim3(64), 0, b->ctx->stream, x.acc);
  AFQ_CHECK_LAUNCH();
  return AFQ_OK;
}

int afq_batch_destroy(afq_batch* b) {
  if (!b) return AFQ_OK;
  if (b->origin) (void)hipStreamSynchronize(b->origin);
  for (afq_ctx* c : b->sctx) (void)hipStreamSynchronize(c->stream);
  if (b->exec[0]) (void)hipGraphExecDestroy(b->exec[0]);
  if (b->exec[1] && b->exec[1] != b->exec[0]) (void)hipGraphExecDestroy(b->exec[1]);
  for (void* p : b->owned) (void)hipFree(p);
  for (hipEvent_t e : b->ev_join) (void)hipEventDestroy(e);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  for (afq_ctx* c : b->sctx) {
    hipStream_t st = c->stream;
    c->comb_static = nullptr;
    afq_ctx_destroy(c);
    (void)hipStreamDestroy(st);
  }
  if (b->origin) (void)hipStreamDestroy(b->origin);
  delete b;
  return AFQ_OK;
}

}  // extern "C"
