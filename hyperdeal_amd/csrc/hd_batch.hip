// hd_batch.hip -- batches of independent advection operators (C ABI):
// one launch per degree and apply / LSRK stage for all of them, an
// autotuner for the task size, and hipGraph capture of a whole LSRK step.
#include "../../include/hyperdeal_amd.h"
#include "hd_device.h"
#include "hd_handles.h"

#include <algorithm>
#include <cmath>
#include <new>
#include <string>
#include <vector>

using hd_abi::as_stream;
using hd_abi::fail;
using hd_abi::hip_fail;
using hd_abi::ok;

namespace
{
constexpr int kMaxN = 8;

int prec_index(int precision)
{
  return precision == HD_FP64 ? 0 : 1;
}
} // namespace

// the work splits of one precision on the device: per degree N present, a
// task list and the blocks' task ranges, all in one allocation (a graph
// keeps its own copy)
struct DevPlan
{
  void *mem = nullptr;
  size_t bytes = 0;
  const hd::BatchTask *tasks[kMaxN + 1] = {};
  hd::BatchPlan plan[kMaxN + 1] = {};
};

struct hd_batch
{
  std::vector<hd_advection *> ops;
  std::vector<long long> n_dofs;
  std::vector<int> degrees; // the N (= degree + 1) present, ascending
  int D = 0;
  hd::BatchOp *dev_ops[2] = {nullptr, nullptr}; // FP64 / FP32 coefficients
  // per precision and N: the measured cost of one DoF (seconds, from the
  // autotuner; 1 before it ran) and the most units per task (0: no limit)
  double weight[2][kMaxN + 1];
  int cap[2][kMaxN + 1];
  DevPlan plan[2];
  std::vector<hd::BatchTask> host_tasks[2][kMaxN + 1];
  bool dirty[2] = {true, true};
};

struct hd_graph
{
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void *plan = nullptr; // the graph's own copy of the work splits
};

namespace
{
// The work split of the operators of degree N (N = degree + 1) over 8 x per
// blocks of that degree's kernel, the single-operator kernel's split for
// every operator: each operator's logical units are cut into eight
// contiguous parts, part x on XCD x (so every XCD gets an eighth of every
// operator: balanced, and its blocks work on neighbouring cells whose faces
// its L2 holds).  An XCD's blocks walk the concatenation of its parts
// interleaved, unit u of the list on block u mod per, and each block
// restarts its load pipeline once per part.  Tasks hold at most cap units
// when cap > 0.
void make_tasks(const hd_batch *b, int pi, int N, int per, int cap,
                std::vector<hd::BatchTask> &tasks,
                std::vector<int> &block_start)
{
  const int cpb = hd::collapsed_batch_cells_per_block(b->D, N, pi);
  tasks.clear();
  block_start.assign(8 * (size_t)per + 1, 0);
  for (int x = 0; x < 8; ++x)
    for (int j = 0; j < per; ++j)
      {
        block_start[(size_t)x * per + j] = (int)tasks.size();
        long long offset = 0; // units of the XCD's list before this part
        for (int i = 0; i < (int)b->ops.size(); ++i)
          {
            const hd_matrix_free *mf = b->ops[i]->mf;
            if (mf->nd != N)
              continue;
            const long long nl = (mf->mesh.n_cells + cpb - 1) / cpb;
            const long long lb = nl * x / 8, le = nl * (x + 1) / 8;
            const long long n = le - lb;
            if (n == 0)
              continue;
            const long long k0 = ((j - offset) % per + per) % per;
            offset += n;
            if (k0 >= n)
              continue;
            const long long n_mine = (n - 1 - k0) / per + 1;
            const long long piece = cap > 0 ? cap : n_mine;
            for (long long m = 0; m < n_mine; m += piece)
              {
                hd::BatchTask task{};
                task.op = i;
                task.stride = per;
                task.lb = lb + k0 + m * per;
                task.le = std::min(le, task.lb + piece * per);
                tasks.push_back(task);
              }
          }
      }
  block_start[8 * (size_t)per] = (int)tasks.size();
}

void free_plan(DevPlan &p)
{
  (void)hipFree(p.mem);
  p = DevPlan{};
}

// the splits of every degree present (caps from b->cap unless `caps`)
int build_plans(hd_batch *b, int pi, const int *caps, DevPlan &p,
                std::vector<hd::BatchTask> *host_tasks)
{
  free_plan(p);
  std::vector<hd::BatchTask> tasks[kMaxN + 1];
  std::vector<int> starts[kMaxN + 1];
  int per[kMaxN + 1] = {};
  size_t bytes = 0;
  for (int N : b->degrees)
    {
      per[N] = hd::collapsed_batch_blocks_per_xcd(N, pi);
      make_tasks(b, pi, N, per[N], caps ? caps[N] : b->cap[pi][N], tasks[N],
                 starts[N]);
      bytes += (tasks[N].size() * sizeof(hd::BatchTask) + 15) / 16 * 16;
      bytes += (starts[N].size() * sizeof(int) + 15) / 16 * 16;
    }
  hipError_t e = hipMalloc(&p.mem, std::max<size_t>(bytes, 16));
  size_t off = 0;
  char *base = static_cast<char *>(p.mem);
  for (int N : b->degrees)
    {
      if (e != hipSuccess)
        break;
      const size_t tb = tasks[N].size() * sizeof(hd::BatchTask);
      if (tb)
        e = hipMemcpy(base + off, tasks[N].data(), tb, hipMemcpyHostToDevice);
      p.tasks[N] = reinterpret_cast<const hd::BatchTask *>(base + off);
      off += (tb + 15) / 16 * 16;
      const size_t sb = starts[N].size() * sizeof(int);
      if (e == hipSuccess)
        e = hipMemcpy(base + off, starts[N].data(), sb, hipMemcpyHostToDevice);
      p.plan[N].block_start = reinterpret_cast<const int *>(base + off);
      p.plan[N].per = per[N];
      off += (sb + 15) / 16 * 16;
      if (host_tasks)
        host_tasks[N] = tasks[N];
    }
  if (e != hipSuccess)
    {
      free_plan(p);
      return hip_fail(e, "batch work split");
    }
  p.bytes = bytes;
  return ok();
}

int ensure_tasks(hd_batch *b, int pi)
{
  if (!b->dirty[pi])
    return ok();
  const int rc = build_plans(b, pi, nullptr, b->plan[pi], b->host_tasks[pi]);
  if (rc != HD_OK)
    return rc;
  b->dirty[pi] = false;
  return ok();
}

// the same splits at another address (a graph's copy): pointers rebased
DevPlan rebase(const hd_batch *b, const DevPlan &p, void *mem)
{
  DevPlan q = p;
  q.mem = mem;
  auto move = [&](const void *ptr) {
    return static_cast<char *>(mem) +
           (static_cast<const char *>(ptr) - static_cast<const char *>(p.mem));
  };
  for (int N : b->degrees)
    {
      q.tasks[N] = reinterpret_cast<const hd::BatchTask *>(move(p.tasks[N]));
      q.plan[N].block_start =
        reinterpret_cast<const int *>(move(p.plan[N].block_start));
    }
  return q;
}

// fn(N, s) for every degree present, in order on `stream`.  (Forking the
// degrees onto concurrent streams measured slower on MI355X: 136-140 against
// 143 GDoF/s for the 64-operator FP32-mixed step -- concurrent kernels no
// longer get the blocks-to-XCD mapping the work split assumes.)
template <class F>
hipError_t for_degrees(hd_batch *b, hipStream_t stream, F fn)
{
  hipError_t e = hipSuccess;
  for (int N : b->degrees)
    if (e == hipSuccess)
      e = fn(N, stream);
  return e;
}

bool step_vectors_ok(const hd_batch *b, double *const *sol,
                     void *const *ti_a, void *const *ti_b)
{
  for (size_t i = 0; i < b->ops.size(); ++i)
    if (ti_a[i] == ti_b[i] || reinterpret_cast<uintptr_t>(sol[i]) % 16 ||
        reinterpret_cast<uintptr_t>(ti_a[i]) % 16 ||
        reinterpret_cast<uintptr_t>(ti_b[i]) % 16)
      return false;
  return true;
}

bool valid_ptrs(const hd_batch *b, const void *const *p)
{
  if (!p)
    return false;
  for (size_t i = 0; i < b->ops.size(); ++i)
    if (!p[i])
      return false;
  return true;
}

// one LSRK step: the FP64 -> precision copy of every solution, then per
// degree the stages of that degree's operators
int batch_step(hd_batch *b, const char *type, int precision,
               double *const *sol, void *const *ti_a, void *const *ti_b,
               double dt, const DevPlan &plan, hipStream_t stream)
{
  double a[16], bb[16];
  const int n_stages = hd_lsrk_coefficients(type, a, bb);
  if (n_stages < 0)
    return n_stages;
  const int n = (int)b->ops.size();
  const int pi = prec_index(precision);
  hd::BatchPtrs p{};
  for (int i = 0; i < n; ++i)
    {
      p.sol[i] = sol[i];
      p.dst[i] = ti_a[i];
    }
  hipError_t e = hd::batch_convert(pi, n, b->n_dofs.data(), p, stream);
  if (e != hipSuccess)
    return hip_fail(e, "batch copy");
  e = for_degrees(b, stream, [&](int N, hipStream_t s) {
    hd::BatchPtrs q = p;
    bool in_a = true;
    hipError_t err = hipSuccess;
    for (int st = 0; st < n_stages && err == hipSuccess; ++st)
      {
        const bool last = st == n_stages - 1;
        for (int i = 0; i < n; ++i)
          {
            q.src[i] = in_a ? ti_a[i] : ti_b[i];
            q.dst[i] = nullptr;
            q.ti[i] = last ? nullptr : (in_a ? ti_b[i] : ti_a[i]);
          }
        q.a_dt = last ? 0.0 : a[st] * dt;
        q.b_dt = bb[st] * dt;
        err = hd::collapsed_batch(b->D, N, pi, last ? 2 : 1, b->dev_ops[pi],
                                  plan.tasks[N], plan.plan[N], q, s);
        in_a = !in_a;
      }
    return err;
  });
  return e == hipSuccess ? ok() : hip_fail(e, "batch stage");
}
} // namespace

extern "C" {

int hd_batch_create(hd_advection *const *ops, int n_ops, hd_batch **out)
{
  if (!ops || !out || n_ops < 1 || n_ops > HD_BATCH_MAX)
    return fail(HD_ERR_INVALID, "hd_batch_create: 1 to " +
                                  std::to_string(HD_BATCH_MAX) +
                                  " operators");
  *out = nullptr;
  const int D = ops[0] ? ops[0]->mf->D : 0;
  for (int i = 0; i < n_ops; ++i)
    {
      const hd_advection *op = ops[i];
      if (!op)
        return fail(HD_ERR_INVALID, "hd_batch_create: null operator");
      if (op->vp || !op->collapsed_ok || op->dirichlet || op->mf->D != D ||
          !hd::collapsed_batch_supported(D, op->mf->nd))
        return fail(HD_ERR_UNSUPPORTED,
                    "batched operators: constant-velocity periodic 2+2 "
                    "meshes, degree <= 5, all of one dimension");
    }
  auto *b = new (std::nothrow) hd_batch();
  if (!b)
    return fail(HD_ERR_NOMEM, "out of host memory");
  b->D = D;
  for (int p = 0; p < 2; ++p)
    for (int N = 0; N <= kMaxN; ++N)
      {
        b->weight[p][N] = 1.0;
        b->cap[p][N] = 0;
      }
  std::vector<hd::BatchOp> table[2];
  for (int i = 0; i < n_ops; ++i)
    {
      b->ops.push_back(ops[i]);
      b->n_dofs.push_back(ops[i]->mf->n_dofs);
      const int N = ops[i]->mf->nd;
      if (std::find(b->degrees.begin(), b->degrees.end(), N) ==
          b->degrees.end())
        b->degrees.push_back(N);
      for (int p = 0; p < 2; ++p)
        {
          hd::BatchOp o{};
          o.coef = p == 0 ? (const void *)ops[i]->coef64
                          : (const void *)ops[i]->coef32;
          o.mesh = ops[i]->mf->mesh;
          o.inflow = ops[i]->inflow;
          o.N = N;
          o.n_dofs = ops[i]->mf->n_dofs;
          table[p].push_back(o);
        }
    }
  std::sort(b->degrees.begin(), b->degrees.end());
  hipError_t e = hipSuccess;
  for (int p = 0; p < 2 && e == hipSuccess; ++p)
    {
      e = hipMalloc(&b->dev_ops[p], n_ops * sizeof(hd::BatchOp));
      if (e == hipSuccess)
        e = hipMemcpy(b->dev_ops[p], table[p].data(),
                      n_ops * sizeof(hd::BatchOp), hipMemcpyHostToDevice);
    }
  if (e != hipSuccess)
    {
      hd_batch_destroy(b);
      return hip_fail(e, "hd_batch_create");
    }
  *out = b;
  return ok();
}

int hd_batch_destroy(hd_batch *batch)
{
  if (!batch)
    return ok();
  for (int p = 0; p < 2; ++p)
    {
      (void)hipFree(batch->dev_ops[p]);
      free_plan(batch->plan[p]);
    }
  delete batch;
  return ok();
}

int hd_batch_size(const hd_batch *batch)
{
  return batch ? (int)batch->ops.size() : HD_ERR_INVALID;
}

int hd_batch_apply(hd_batch *batch, const void *const *src, void *const *dst,
                   int precision, void *stream)
{
  if (!batch || !valid_ptrs(batch, src) ||
      !valid_ptrs(batch, (const void *const *)dst) ||
      (precision != HD_FP64 && precision != HD_FP32))
    return fail(HD_ERR_INVALID, "hd_batch_apply: invalid argument");
  const int pi = prec_index(precision);
  int rc = ensure_tasks(batch, pi);
  if (rc != HD_OK)
    return rc;
  hd::BatchPtrs p{};
  for (size_t i = 0; i < batch->ops.size(); ++i)
    {
      if (reinterpret_cast<uintptr_t>(src[i]) % 16 != 0)
        return fail(HD_ERR_INVALID,
                    "hd_batch_apply: source vectors must be 16-byte aligned");
      p.src[i] = src[i];
      p.dst[i] = dst[i];
    }
  const DevPlan &pl = batch->plan[pi];
  const hipError_t e =
    for_degrees(batch, as_stream(stream), [&](int N, hipStream_t s) {
      return hd::collapsed_batch(batch->D, N, pi, 0, batch->dev_ops[pi],
                                 pl.tasks[N], pl.plan[N], p, s);
    });
  return e == hipSuccess ? ok() : hip_fail(e, "hd_batch_apply");
}

int hd_batch_lsrk_step(hd_batch *batch, const char *type, int precision,
                       double *const *sol, void *const *ti_a,
                       void *const *ti_b, double dt, void *stream)
{
  if (!batch || !valid_ptrs(batch, (const void *const *)sol) ||
      !valid_ptrs(batch, (const void *const *)ti_a) ||
      !valid_ptrs(batch, (const void *const *)ti_b) ||
      (precision != HD_FP64 && precision != HD_FP32))
    return fail(HD_ERR_INVALID, "hd_batch_lsrk_step: invalid argument");
  if (!step_vectors_ok(batch, sol, ti_a, ti_b))
    return fail(HD_ERR_INVALID, "hd_batch_lsrk_step: the two stage "
                                "registers must differ; sol and both "
                                "registers must be 16-byte aligned");
  const int pi = prec_index(precision);
  const int rc = ensure_tasks(batch, pi);
  if (rc != HD_OK)
    return rc;
  return batch_step(batch, type, precision, sol, ti_a, ti_b, dt,
                    batch->plan[pi], as_stream(stream));
}

int hd_batch_autotune(hd_batch *batch, int precision, const void *const *src,
                      void *const *dst, void *stream)
{
  if (!batch || !valid_ptrs(batch, src) ||
      !valid_ptrs(batch, (const void *const *)dst) ||
      (precision != HD_FP64 && precision != HD_FP32))
    return fail(HD_ERR_INVALID, "hd_batch_autotune: invalid argument");
  const int pi = prec_index(precision);
  hipStream_t s = as_stream(stream);
  hd::BatchPtrs p{};
  for (size_t i = 0; i < batch->ops.size(); ++i)
    {
      p.src[i] = src[i];
      p.dst[i] = dst[i];
    }
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipError_t e = hipEventCreate(&t0);
  if (e == hipSuccess)
    e = hipEventCreate(&t1);
  if (e != hipSuccess)
    return hip_fail(e, "hd_batch_autotune");
  // per degree, the task size (units per task; 0: a block's whole share of
  // a fragment) that applies that degree's operators fastest: one warm-up,
  // then the mean of three applies
  constexpr int kCaps[] = {0, 1, 4, 16};
  int best_cap[kMaxN + 1] = {};
  int rc = ok();
  for (int N : batch->degrees)
    {
      float best = -1.0f;
      double dofs = 0.0;
      for (size_t i = 0; i < batch->ops.size(); ++i)
        if (batch->ops[i]->mf->nd == N)
          dofs += (double)batch->ops[i]->mf->n_dofs;
      for (int cap : kCaps)
        {
          if (rc != HD_OK)
            break;
          int caps[kMaxN + 1] = {};
          caps[N] = cap;
          DevPlan dp;
          rc = build_plans(batch, pi, caps, dp, nullptr);
          if (rc != HD_OK)
            break;
          hipError_t err = hd::collapsed_batch(batch->D, N, pi, 0,
                                               batch->dev_ops[pi],
                                               dp.tasks[N], dp.plan[N], p, s);
          if (err == hipSuccess)
            err = hipEventRecord(t0, s);
          for (int r = 0; r < 3 && err == hipSuccess; ++r)
            err = hd::collapsed_batch(batch->D, N, pi, 0, batch->dev_ops[pi],
                                      dp.tasks[N], dp.plan[N], p, s);
          if (err == hipSuccess)
            err = hipEventRecord(t1, s);
          if (err == hipSuccess)
            err = hipEventSynchronize(t1);
          float ms = 0.0f;
          if (err == hipSuccess)
            err = hipEventElapsedTime(&ms, t0, t1);
          free_plan(dp);
          if (err != hipSuccess)
            {
              rc = hip_fail(err, "hd_batch_autotune");
              break;
            }
          ms /= 3.0f;
          if (best < 0.0f || ms < best)
            {
              best = ms;
              best_cap[N] = cap;
            }
        }
      if (rc == HD_OK)
        batch->weight[pi][N] = std::max(1e-3 * best / dofs, 1e-18);
    }
  if (rc == HD_OK)
    for (int N : batch->degrees)
      batch->cap[pi][N] = best_cap[N];
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  batch->dirty[pi] = true;
  return rc;
}

int hd_batch_cells_per_task(const hd_batch *batch, int degree, int precision)
{
  if (!batch || degree < 1 || degree + 1 > kMaxN ||
      (precision != HD_FP64 && precision != HD_FP32))
    return fail(HD_ERR_INVALID, "hd_batch_cells_per_task: invalid argument");
  const int pi = prec_index(precision);
  const int N = degree + 1;
  const int cpb = hd::collapsed_batch_cells_per_block(batch->D, N, pi);
  if (cpb <= 0)
    return fail(HD_ERR_UNSUPPORTED, "no batch kernel for this degree");
  if (std::find(batch->degrees.begin(), batch->degrees.end(), N) ==
      batch->degrees.end())
    return (batch->cap[pi][N] > 0 ? batch->cap[pi][N] : 1) * cpb;
  std::vector<hd::BatchTask> fresh;
  const std::vector<hd::BatchTask> *tasks = &batch->host_tasks[pi][N];
  if (batch->dirty[pi])
    {
      std::vector<int> block_start;
      make_tasks(batch, pi, N, hd::collapsed_batch_blocks_per_xcd(N, pi),
                 batch->cap[pi][N], fresh, block_start);
      tasks = &fresh;
    }
  long long units = 0, count = 0;
  for (const hd::BatchTask &t : *tasks)
    {
      units += (t.le - t.lb + t.stride - 1) / t.stride;
      ++count;
    }
  if (count == 0)
    return (batch->cap[pi][N] > 0 ? batch->cap[pi][N] : 1) * cpb;
  return (int)std::max(1LL, units * cpb / count);
}

int hd_batch_set_cells_per_task(hd_batch *batch, int degree, int precision,
                                int cells)
{
  if (!batch || degree < 1 || degree + 1 > kMaxN || cells < 1 ||
      (precision != HD_FP64 && precision != HD_FP32))
    return fail(HD_ERR_INVALID,
                "hd_batch_set_cells_per_task: invalid argument");
  const int pi = prec_index(precision);
  const int N = degree + 1;
  const int cpb = hd::collapsed_batch_cells_per_block(batch->D, N, pi);
  if (cpb <= 0)
    return fail(HD_ERR_UNSUPPORTED, "no batch kernel for this degree");
  batch->cap[pi][N] = std::max(1, (cells + cpb - 1) / cpb);
  batch->dirty[pi] = true;
  return ok();
}

int hd_batch_graph_create(hd_batch *batch, const char *type, int precision,
                          double *const *sol, void *const *ti_a,
                          void *const *ti_b, double dt, hd_graph **out)
{
  if (!batch || !out || !valid_ptrs(batch, (const void *const *)sol) ||
      !valid_ptrs(batch, (const void *const *)ti_a) ||
      !valid_ptrs(batch, (const void *const *)ti_b) ||
      (precision != HD_FP64 && precision != HD_FP32))
    return fail(HD_ERR_INVALID, "hd_batch_graph_create: invalid argument");
  *out = nullptr;
  if (!step_vectors_ok(batch, sol, ti_a, ti_b))
    return fail(HD_ERR_INVALID, "hd_batch_graph_create: the two stage "
                                "registers must differ; sol and both "
                                "registers must be 16-byte aligned");
  const int pi = prec_index(precision);
  int rc = ensure_tasks(batch, pi);
  if (rc != HD_OK)
    return rc;
  auto *g = new (std::nothrow) hd_graph();
  if (!g)
    return fail(HD_ERR_NOMEM, "out of host memory");
  const DevPlan &bp = batch->plan[pi];
  hipError_t e = hipMalloc(&g->plan, std::max<size_t>(bp.bytes, 16));
  if (e == hipSuccess && bp.bytes)
    e = hipMemcpy(g->plan, bp.mem, bp.bytes, hipMemcpyDeviceToDevice);
  const DevPlan gp = rebase(batch, bp, g->plan);
  hipStream_t s = nullptr;
  if (e == hipSuccess)
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e != hipSuccess)
    {
      hd_graph_destroy(g);
      return hip_fail(e, "hd_batch_graph_create");
    }
  e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess)
    {
      rc = batch_step(batch, type, precision, sol, ti_a, ti_b, dt, gp, s);
      const hipError_t e2 = hipStreamEndCapture(s, &g->graph);
      if (rc == HD_OK && e2 != hipSuccess)
        rc = hip_fail(e2, "hd_batch_graph_create: capture");
    }
  else
    rc = hip_fail(e, "hd_batch_graph_create: capture");
  if (rc == HD_OK)
    {
      e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
      if (e != hipSuccess)
        rc = hip_fail(e, "hd_batch_graph_create: instantiate");
    }
  (void)hipStreamDestroy(s);
  if (rc != HD_OK)
    {
      hd_graph_destroy(g);
      return rc;
    }
  *out = g;
  return ok();
}

int hd_graph_launch(hd_graph *graph, void *stream)
{
  if (!graph || !graph->exec)
    return fail(HD_ERR_INVALID, "hd_graph_launch: invalid graph");
  const hipError_t e = hipGraphLaunch(graph->exec, as_stream(stream));
  return e == hipSuccess ? ok() : hip_fail(e, "hd_graph_launch");
}

int hd_graph_destroy(hd_graph *graph)
{
  if (!graph)
    return ok();
  if (graph->exec)
    (void)hipGraphExecDestroy(graph->exec);
  if (graph->graph)
    (void)hipGraphDestroy(graph->graph);
  (void)hipFree(graph->plan);
  delete graph;
  return ok();
}

} // extern "C"
