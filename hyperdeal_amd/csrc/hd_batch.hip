// hd_batch.hip -- batches of independent advection operators (C ABI):
// one launch per apply / LSRK stage for all of them, an autotuner for the
// work split, and hipGraph capture of a whole LSRK step.
#include "../../include/hyperdeal_amd.h"
#include "hd_device.h"
#include "hd_handles.h"

#include <algorithm>
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
// candidate work items per task, in units of the cells one block handles at a
// time; the autotuner times each
constexpr int kCandidates[] = {1, 2, 4, 8, 16, 32};
constexpr int kDefaultUnits = 4;

int prec_index(int precision)
{
  return precision == HD_FP64 ? 0 : 1;
}
} // namespace

struct hd_batch
{
  std::vector<hd_advection *> ops;
  std::vector<long long> n_dofs;
  int D = 0;
  hd::BatchOp *dev_ops[2] = {nullptr, nullptr}; // FP64 / FP32 coefficients
  int units[2][kMaxN + 1];                      // units per task, per N
  hd::BatchTask *dev_tasks[2] = {nullptr, nullptr};
  hd::BatchSegments seg[2] = {};
  bool dirty[2] = {true, true};
};

struct hd_graph
{
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hd::BatchTask *tasks = nullptr; // the graph's own copy of the task list
};

namespace
{
// tasks of the operators in `which` (all if empty) in one segment per XCD:
// each operator's logical units are cut into 8 contiguous ranges, range x
// goes to XCD x (so the face neighbours of a range's cells sit in that XCD's
// L2, as in the single-operator kernel) in tasks of units[N] units; each
// segment runs its longest tasks first
void make_tasks(const hd_batch *b, int pi, const std::vector<int> &which,
                const int *units, std::vector<hd::BatchTask> &tasks,
                hd::BatchSegments &seg)
{
  std::vector<std::pair<double, hd::BatchTask>> t[8];
  for (int i = 0; i < (int)b->ops.size(); ++i)
    {
      if (!which.empty() &&
          std::find(which.begin(), which.end(), i) == which.end())
        continue;
      const hd_matrix_free *mf = b->ops[i]->mf;
      const int N = mf->nd;
      const int cpb = hd::collapsed_batch_cells_per_block(b->D, N, pi);
      const long long n_logical = (mf->mesh.n_cells + cpb - 1) / cpb;
      const int u = units[N];
      for (int x = 0; x < 8; ++x)
        {
          const long long xb = n_logical * x / 8, xe = n_logical * (x + 1) / 8;
          for (long long lb = xb; lb < xe; lb += u)
            {
              hd::BatchTask task{};
              task.op = i;
              task.lb = lb;
              task.le = std::min(lb + u, xe);
              // cost: the DoF the task touches
              const double cost =
                (double)(task.le - task.lb) * cpb * mf->dofs_cell;
              t[x].push_back({cost, task});
            }
        }
    }
  tasks.clear();
  for (int x = 0; x < 8; ++x)
    {
      std::stable_sort(t[x].begin(), t[x].end(),
                       [](const auto &p, const auto &q) {
                         return p.first > q.first;
                       });
      seg.start[x] = (int)tasks.size();
      for (const auto &y : t[x])
        tasks.push_back(y.second);
    }
  seg.start[8] = (int)tasks.size();
}

int upload_tasks(const std::vector<hd::BatchTask> &tasks,
                 hd::BatchTask **dev)
{
  *dev = nullptr;
  if (tasks.empty())
    return ok();
  hipError_t e = hipMalloc(dev, tasks.size() * sizeof(hd::BatchTask));
  if (e == hipSuccess)
    e = hipMemcpy(*dev, tasks.data(), tasks.size() * sizeof(hd::BatchTask),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess)
    {
      (void)hipFree(*dev);
      *dev = nullptr;
      return hip_fail(e, "batch task list");
    }
  return ok();
}

int ensure_tasks(hd_batch *b, int pi)
{
  if (!b->dirty[pi])
    return ok();
  std::vector<hd::BatchTask> tasks;
  hd::BatchSegments seg{};
  make_tasks(b, pi, {}, b->units[pi], tasks, seg);
  if (b->dev_tasks[pi])
    (void)hipFree(b->dev_tasks[pi]);
  b->dev_tasks[pi] = nullptr;
  const int rc = upload_tasks(tasks, &b->dev_tasks[pi]);
  if (rc != HD_OK)
    return rc;
  b->seg[pi] = seg;
  b->dirty[pi] = false;
  return ok();
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

// the step: copy sol into ti_a, then the stages ping-ponging ti_a / ti_b
int batch_step(hd_batch *b, const char *type, int precision,
               double *const *sol, void *const *ti_a, void *const *ti_b,
               double dt, const hd::BatchTask *tasks,
               const hd::BatchSegments &seg, hipStream_t stream)
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
  // the first stage reads a copy of the solution (time_integrators.
  // templates.h:142-157)
  hipError_t e =
    hd::batch_convert(pi, n, b->n_dofs.data(), p, stream);
  if (e != hipSuccess)
    return hip_fail(e, "batch copy");
  bool in_a = true;
  for (int s = 0; s < n_stages; ++s)
    {
      const bool last = s == n_stages - 1;
      for (int i = 0; i < n; ++i)
        {
          p.src[i] = in_a ? ti_a[i] : ti_b[i];
          p.dst[i] = nullptr;
          p.ti[i] = last ? nullptr : (in_a ? ti_b[i] : ti_a[i]);
        }
      p.a_dt = last ? 0.0 : a[s] * dt;
      p.b_dt = bb[s] * dt;
      e = hd::collapsed_batch(b->D, pi, last ? 2 : 1, b->dev_ops[pi], tasks,
                              seg, p, stream);
      if (e != hipSuccess)
        return hip_fail(e, "batch stage");
      in_a = !in_a;
    }
  return ok();
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
      if (op->vp || !op->collapsed_ok || op->mf->D != D ||
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
      b->units[p][N] = kDefaultUnits;
  std::vector<hd::BatchOp> table[2];
  for (int i = 0; i < n_ops; ++i)
    {
      b->ops.push_back(ops[i]);
      b->n_dofs.push_back(ops[i]->mf->n_dofs);
      for (int p = 0; p < 2; ++p)
        {
          hd::BatchOp o{};
          o.coef = p == 0 ? (const void *)ops[i]->coef64
                          : (const void *)ops[i]->coef32;
          o.mesh = ops[i]->mf->mesh;
          o.inflow = ops[i]->inflow;
          o.N = ops[i]->mf->nd;
          o.n_dofs = ops[i]->mf->n_dofs;
          table[p].push_back(o);
        }
    }
  for (int p = 0; p < 2; ++p)
    {
      hipError_t e = hipMalloc(&b->dev_ops[p], n_ops * sizeof(hd::BatchOp));
      if (e == hipSuccess)
        e = hipMemcpy(b->dev_ops[p], table[p].data(),
                      n_ops * sizeof(hd::BatchOp), hipMemcpyHostToDevice);
      if (e != hipSuccess)
        {
          hd_batch_destroy(b);
          return hip_fail(e, "hd_batch_create");
        }
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
      (void)hipFree(batch->dev_tasks[p]);
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
  const hipError_t e =
    hd::collapsed_batch(batch->D, pi, 0, batch->dev_ops[pi],
                        batch->dev_tasks[pi], batch->seg[pi], p,
                        as_stream(stream));
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
  for (size_t i = 0; i < batch->ops.size(); ++i)
    if (ti_a[i] == ti_b[i] || reinterpret_cast<uintptr_t>(ti_a[i]) % 16 ||
        reinterpret_cast<uintptr_t>(ti_b[i]) % 16)
      return fail(HD_ERR_INVALID, "hd_batch_lsrk_step: the two stage "
                                  "registers must differ and be 16-byte "
                                  "aligned");
  const int pi = prec_index(precision);
  const int rc = ensure_tasks(batch, pi);
  if (rc != HD_OK)
    return rc;
  return batch_step(batch, type, precision, sol, ti_a, ti_b, dt,
                    batch->dev_tasks[pi], batch->seg[pi], as_stream(stream));
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
  int rc = ok();
  // per degree: the operators of that degree alone, each candidate timed
  for (int N = 2; N <= kMaxN && rc == HD_OK; ++N)
    {
      std::vector<int> which;
      for (size_t i = 0; i < batch->ops.size(); ++i)
        if (batch->ops[i]->mf->nd == N)
          which.push_back((int)i);
      if (which.empty())
        continue;
      float best = 1e30f;
      int best_u = kDefaultUnits;
      for (int u : kCandidates)
        {
          int units[kMaxN + 1];
          std::fill(units, units + kMaxN + 1, u);
          std::vector<hd::BatchTask> tasks;
          hd::BatchSegments seg{};
          make_tasks(batch, pi, which, units, tasks, seg);
          hd::BatchTask *dev = nullptr;
          rc = upload_tasks(tasks, &dev);
          if (rc != HD_OK)
            break;
          float ms = 0.0f;
          e = hd::collapsed_batch(batch->D, pi, 0, batch->dev_ops[pi], dev,
                                  seg, p, s); // warm-up
          if (e == hipSuccess)
            e = hipEventRecord(t0, s);
          for (int r = 0; r < 3 && e == hipSuccess; ++r)
            e = hd::collapsed_batch(batch->D, pi, 0, batch->dev_ops[pi], dev,
                                    seg, p, s);
          if (e == hipSuccess)
            e = hipEventRecord(t1, s);
          if (e == hipSuccess)
            e = hipEventSynchronize(t1);
          if (e == hipSuccess)
            e = hipEventElapsedTime(&ms, t0, t1);
          (void)hipFree(dev);
          if (e != hipSuccess)
            {
              rc = hip_fail(e, "hd_batch_autotune");
              break;
            }
          if (ms < best)
            {
              best = ms;
              best_u = u;
            }
        }
      if (rc == HD_OK)
        batch->units[pi][N] = best_u;
    }
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
  return batch->units[pi][N] * cpb;
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
  batch->units[pi][N] = std::max(1, (cells + cpb - 1) / cpb);
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
  for (size_t i = 0; i < batch->ops.size(); ++i)
    if (ti_a[i] == ti_b[i])
      return fail(HD_ERR_INVALID, "hd_batch_graph_create: the two stage "
                                  "registers must differ");
  const int pi = prec_index(precision);
  int rc = ensure_tasks(batch, pi);
  if (rc != HD_OK)
    return rc;
  auto *g = new (std::nothrow) hd_graph();
  if (!g)
    return fail(HD_ERR_NOMEM, "out of host memory");
  // the graph keeps its own task list: later tuning does not change it
  const int n_tasks = batch->seg[pi].start[8];
  hipError_t e = hipMalloc(&g->tasks, std::max(n_tasks, 1) *
                                        sizeof(hd::BatchTask));
  if (e == hipSuccess && n_tasks > 0)
    e = hipMemcpy(g->tasks, batch->dev_tasks[pi],
                  n_tasks * sizeof(hd::BatchTask), hipMemcpyDeviceToDevice);
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
      rc = batch_step(batch, type, precision, sol, ti_a, ti_b, dt, g->tasks,
                      batch->seg[pi], s);
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
  (void)hipFree(graph->tasks);
  delete graph;
  return ok();
}

} // extern "C"
