// hd_batch.hip -- batches of independent advection operators (C ABI):
// one launch per apply / LSRK stage for all of them, an autotuner for the
// work split, and hipGraph capture of a whole LSRK step.
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
// the autotuner's refinement: factors tried on each degree's cost weight
constexpr double kRefine[] = {0.8, 1.25};

int prec_index(int precision)
{
  return precision == HD_FP64 ? 0 : 1;
}
} // namespace

// a work split on the device: the task list, then the blocks' task ranges
struct DevPlan
{
  void *mem = nullptr;
  size_t bytes = 0;
  const hd::BatchTask *tasks = nullptr;
  hd::BatchPlan plan{};
};

struct hd_batch
{
  std::vector<hd_advection *> ops;
  std::vector<long long> n_dofs;
  int D = 0;
  hd::BatchOp *dev_ops[2] = {nullptr, nullptr}; // FP64 / FP32 coefficients
  // per precision and N: the cost of one DoF in the work split (seconds per
  // DoF from the autotuner, 1 before it ran) and the most units per task
  // (0: no limit)
  double weight[2][kMaxN + 1];
  int cap[2][kMaxN + 1];
  DevPlan plan[2];
  std::vector<hd::BatchTask> host_tasks[2];
  bool dirty[2] = {true, true};
};

struct hd_graph
{
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void *plan = nullptr; // the graph's own copy of the work split
};

namespace
{
// The work split of the operators in `which` (all if empty) over 8 x per
// blocks.  Operators are cut into fragments of contiguous logical units, no
// more than about a sixteenth of the total cost each (cost: weight[N] per
// DoF), and the fragments go to the XCDs, largest first, each to the least
// loaded one.  An XCD's blocks walk its fragments the way the single-operator
// kernel walks its range: interleaved, unit u of the XCD's concatenated list
// on block u mod per, so that the blocks of an XCD work on a window of
// neighbouring cells (their face neighbours in that XCD's L2) and each block
// restarts its load pipeline once per fragment.  Tasks hold at most cap[N]
// units when a cap is set.
void make_tasks(const hd_batch *b, int pi, const std::vector<int> &which,
                int per, std::vector<hd::BatchTask> &tasks,
                std::vector<int> &block_start)
{
  struct Frag
  {
    int op, N;
    long long lb, le;
    double cost;
  };
  std::vector<Frag> frags;
  double total = 0.0;
  for (int i = 0; i < (int)b->ops.size(); ++i)
    {
      if (!which.empty() &&
          std::find(which.begin(), which.end(), i) == which.end())
        continue;
      const hd_matrix_free *mf = b->ops[i]->mf;
      const int cpb = hd::collapsed_batch_cells_per_block(b->D, mf->nd, pi);
      const long long nl = (mf->mesh.n_cells + cpb - 1) / cpb;
      total += b->weight[pi][mf->nd] * (double)nl * cpb * mf->dofs_cell;
    }
  const double frag_cost = total / 16.0;
  for (int i = 0; i < (int)b->ops.size(); ++i)
    {
      if (!which.empty() &&
          std::find(which.begin(), which.end(), i) == which.end())
        continue;
      const hd_matrix_free *mf = b->ops[i]->mf;
      const int N = mf->nd;
      const int cpb = hd::collapsed_batch_cells_per_block(b->D, N, pi);
      const long long nl = (mf->mesh.n_cells + cpb - 1) / cpb;
      const double uc = b->weight[pi][N] * (double)cpb * mf->dofs_cell;
      const long long nf = std::max(
        1LL, std::min(nl, (long long)std::ceil(nl * uc / frag_cost - 1e-9)));
      for (long long f = 0; f < nf; ++f)
        {
          const long long lb = nl * f / nf, le = nl * (f + 1) / nf;
          frags.push_back({i, N, lb, le, (double)(le - lb) * uc});
        }
    }
  // largest first, each to the least loaded XCD
  std::vector<int> order(frags.size());
  for (size_t f = 0; f < frags.size(); ++f)
    order[f] = (int)f;
  std::stable_sort(order.begin(), order.end(), [&](int p, int q) {
    return frags[p].cost > frags[q].cost;
  });
  std::vector<int> owner(frags.size(), 0);
  double load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int f : order)
    {
      const int x = (int)(std::min_element(load, load + 8) - load);
      owner[f] = x;
      load[x] += frags[f].cost;
    }
  tasks.clear();
  block_start.assign(8 * (size_t)per + 1, 0);
  for (int x = 0; x < 8; ++x)
    {
      // this XCD's fragments by degree, then operator and position
      std::vector<int> mine;
      for (size_t f = 0; f < frags.size(); ++f)
        if (owner[f] == x)
          mine.push_back((int)f);
      std::stable_sort(mine.begin(), mine.end(), [&](int p, int q) {
        return frags[p].N < frags[q].N;
      });
      for (int j = 0; j < per; ++j)
        {
          block_start[(size_t)x * per + j] = (int)tasks.size();
          long long offset = 0; // units of the XCD's list before fragment
          for (int f : mine)
            {
              const Frag &fr = frags[f];
              const long long n = fr.le - fr.lb;
              // the fragment's first unit on block j
              const long long k0 = ((j - offset) % per + per) % per;
              offset += n;
              if (k0 >= n)
                continue;
              const long long n_mine = (n - 1 - k0) / per + 1;
              const int cap = b->cap[pi][fr.N];
              const long long piece = cap > 0 ? cap : n_mine;
              for (long long m = 0; m < n_mine; m += piece)
                {
                  hd::BatchTask task{};
                  task.op = fr.op;
                  task.stride = per;
                  task.lb = fr.lb + k0 + m * per;
                  task.le = std::min(fr.le, task.lb + piece * per);
                  tasks.push_back(task);
                }
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

int upload_plan(const std::vector<hd::BatchTask> &tasks,
                const std::vector<int> &block_start, int per, DevPlan &p)
{
  free_plan(p);
  const size_t tb = (tasks.size() * sizeof(hd::BatchTask) + 15) / 16 * 16;
  p.bytes = tb + block_start.size() * sizeof(int);
  hipError_t e = hipMalloc(&p.mem, p.bytes);
  if (e == hipSuccess && !tasks.empty())
    e = hipMemcpy(p.mem, tasks.data(), tasks.size() * sizeof(hd::BatchTask),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(static_cast<char *>(p.mem) + tb, block_start.data(),
                  block_start.size() * sizeof(int), hipMemcpyHostToDevice);
  if (e != hipSuccess)
    {
      free_plan(p);
      return hip_fail(e, "batch work split");
    }
  p.tasks = static_cast<const hd::BatchTask *>(p.mem);
  p.plan.block_start =
    reinterpret_cast<const int *>(static_cast<char *>(p.mem) + tb);
  p.plan.per = per;
  return ok();
}

int ensure_tasks(hd_batch *b, int pi)
{
  if (!b->dirty[pi])
    return ok();
  const int per = hd::collapsed_batch_blocks_per_xcd(pi);
  std::vector<int> block_start;
  make_tasks(b, pi, {}, per, b->host_tasks[pi], block_start);
  const int rc = upload_plan(b->host_tasks[pi], block_start, per, b->plan[pi]);
  if (rc != HD_OK)
    return rc;
  b->dirty[pi] = false;
  return ok();
}

// the vectors of an LSRK step: distinct stage registers, 16-byte aligned
// (the batched copy and the stage kernels use 16-byte accesses)
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

// the step: copy sol into ti_a, then the stages ping-ponging ti_a / ti_b
int batch_step(hd_batch *b, const char *type, int precision,
               double *const *sol, void *const *ti_a, void *const *ti_b,
               double dt, const hd::BatchTask *tasks,
               const hd::BatchPlan &plan, hipStream_t stream)
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
                              plan, p, stream);
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
      {
        b->weight[p][N] = 1.0;
        b->cap[p][N] = 0;
      }
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
  const hipError_t e =
    hd::collapsed_batch(batch->D, pi, 0, batch->dev_ops[pi],
                        batch->plan[pi].tasks, batch->plan[pi].plan, p,
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
  if (!step_vectors_ok(batch, sol, ti_a, ti_b))
    return fail(HD_ERR_INVALID, "hd_batch_lsrk_step: the two stage "
                                "registers must differ; sol and both "
                                "registers must be 16-byte aligned");
  const int pi = prec_index(precision);
  const int rc = ensure_tasks(batch, pi);
  if (rc != HD_OK)
    return rc;
  return batch_step(batch, type, precision, sol, ti_a, ti_b, dt,
                    batch->plan[pi].tasks, batch->plan[pi].plan,
                    as_stream(stream));
}

int hd_batch_autotune(hd_batch *batch, int precision, const void *const *src,
                      void *const *dst, void *stream)
{
  if (!batch || !valid_ptrs(batch, src) ||
      !valid_ptrs(batch, (const void *const *)dst) ||
      (precision != HD_FP64 && precision != HD_FP32))
    return fail(HD_ERR_INVALID, "hd_batch_autotune: invalid argument");
  const int pi = prec_index(precision);
  const int per = hd::collapsed_batch_blocks_per_xcd(pi);
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
  // milliseconds of one apply of the operators in `which` under the current
  // weights (one warm-up, the mean of three)
  auto time_split = [&](const std::vector<int> &which, float &ms) -> int {
    std::vector<hd::BatchTask> tasks;
    std::vector<int> block_start;
    make_tasks(batch, pi, which, per, tasks, block_start);
    DevPlan dp;
    int rc = upload_plan(tasks, block_start, per, dp);
    if (rc != HD_OK)
      return rc;
    hipError_t err = hd::collapsed_batch(batch->D, pi, 0, batch->dev_ops[pi],
                                         dp.tasks, dp.plan, p, s);
    if (err == hipSuccess)
      err = hipEventRecord(t0, s);
    for (int r = 0; r < 3 && err == hipSuccess; ++r)
      err = hd::collapsed_batch(batch->D, pi, 0, batch->dev_ops[pi], dp.tasks,
                                dp.plan, p, s);
    if (err == hipSuccess)
      err = hipEventRecord(t1, s);
    if (err == hipSuccess)
      err = hipEventSynchronize(t1);
    if (err == hipSuccess)
      err = hipEventElapsedTime(&ms, t0, t1);
    free_plan(dp);
    ms /= 3.0f;
    return err == hipSuccess ? ok() : hip_fail(err, "hd_batch_autotune");
  };
  int rc = ok();
  // 1) each degree alone: its time per DoF is its cost weight
  std::vector<int> degrees;
  for (int N = 2; N <= kMaxN && rc == HD_OK; ++N)
    {
      std::vector<int> which;
      double dofs = 0.0;
      for (size_t i = 0; i < batch->ops.size(); ++i)
        if (batch->ops[i]->mf->nd == N)
          {
            which.push_back((int)i);
            dofs += (double)batch->ops[i]->mf->n_dofs;
          }
      if (which.empty())
        continue;
      float ms = 0.0f;
      rc = time_split(which, ms);
      if (rc == HD_OK)
        {
          batch->weight[pi][N] = std::max(1e-3 * ms / dofs, 1e-18);
          degrees.push_back(N);
        }
    }
  // 2) the whole batch: scale each degree's weight while that helps
  if (rc == HD_OK && degrees.size() > 1)
    {
      float best = 0.0f;
      rc = time_split({}, best);
      for (int N : degrees)
        for (double f : kRefine)
          {
            if (rc != HD_OK)
              break;
            const double w0 = batch->weight[pi][N];
            batch->weight[pi][N] = w0 * f;
            float ms = 0.0f;
            rc = time_split({}, ms);
            if (rc == HD_OK && ms < best)
              best = ms;
            else
              batch->weight[pi][N] = w0;
          }
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
  // the mean task of this degree in the current work split
  std::vector<hd::BatchTask> fresh;
  const std::vector<hd::BatchTask> *tasks = &batch->host_tasks[pi];
  if (batch->dirty[pi])
    {
      std::vector<int> block_start;
      make_tasks(batch, pi, {}, hd::collapsed_batch_blocks_per_xcd(pi), fresh,
                 block_start);
      tasks = &fresh;
    }
  long long units = 0, count = 0;
  for (const hd::BatchTask &t : *tasks)
    if (batch->ops[t.op]->mf->nd == N)
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
  // the graph keeps its own work split: later tuning does not change it
  const DevPlan &bp = batch->plan[pi];
  hipError_t e = hipMalloc(&g->plan, bp.bytes);
  if (e == hipSuccess)
    e = hipMemcpy(g->plan, bp.mem, bp.bytes, hipMemcpyDeviceToDevice);
  const char *base = static_cast<const char *>(g->plan);
  const auto *gtasks = reinterpret_cast<const hd::BatchTask *>(
    base + (reinterpret_cast<const char *>(bp.tasks) -
            static_cast<const char *>(bp.mem)));
  hd::BatchPlan gplan = bp.plan;
  gplan.block_start = reinterpret_cast<const int *>(
    base + (reinterpret_cast<const char *>(bp.plan.block_start) -
            static_cast<const char *>(bp.mem)));
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
      rc = batch_step(batch, type, precision, sol, ti_a, ti_b, dt, gtasks,
                      gplan, s);
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
