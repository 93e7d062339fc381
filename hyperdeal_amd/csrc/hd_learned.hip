// hd_learned.hip -- the C ABI of the convolutional network (hd_cnn_*) and the
// learned initial guess of the Vlasov-Poisson potential solve
// (hd_vp_learned_guess_*).  Neither has a counterpart in the reference,
// whose solve starts from the previous potential
// (vlasov_poisson/include/application.h:575-576); that start stays one of
// the two directions the solver combines, and CG to the relative tolerance
// decides the result, so a poor guess costs one extra matrix-vector product
// and never the accuracy.
//
// Per solve, with the guess on (phi_prev: the potential before the solve):
//   r0     <- b - A phi_prev                    (the start's residual)
//   image  <- (r0 - mean) / rms                 (a ring slot: the training
//                                               input of this solve)
//   g      <- net(image)                        once the net has trained
//   CG from the energy-minimising combination of phi_prev and g
//   target <- (phi - phi_prev - mean) / rms     (the ring slot's target: the
//                                               correction CG found, the
//                                               solution for right-hand
//                                               side r0)
// and, on every solve of the first dense_solves, then every train_every
// solves, train_steps Adam steps on batches drawn on the device from the
// ring's complete pairs.  The training and the target image run on a stream
// of their own, forked after the guess (and after the CG) and joined before
// the next solve: the CG of a solve runs in one workgroup, so the rest of
// the GPU trains meanwhile.
// The network learns the shape of A^-1 on the residuals this run produces;
// the combination sets the scale.  Nothing here synchronises with the host.
#include "hd_handles.h"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>

using hd_abi::fail;
using hd_abi::hip_fail;
using hd_abi::ok;
using hd_abi::as_stream;

namespace hd_abi
{
hipError_t learned_before_solve(hd_advection &op, const double **guess,
                                hipStream_t s)
{
  hd_learned_guess &lg = *op.learned;
  *guess = nullptr;
  hipError_t e = hipSuccess;
  // the training stream's work (the previous target image, the previous
  // training step) is done: the weights are updated, phi_prev is free and
  // no training batch reads the ring slot this solve writes
  if (lg.pending)
    {
      e = hipStreamWaitEvent(s, lg.ev_done, 0);
      lg.pending = false;
    }
  const size_t bytes = (size_t)op.n_x_dofs * sizeof(double);
  if (e == hipSuccess)
    e = hipMemcpyAsync(lg.phi_prev, op.phi, bytes, hipMemcpyDeviceToDevice,
                       s);
  // the residual image straight from the solver's kernel
  float *slot = lg.ring_in + (long long)lg.head * lg.plane;
  hd::ResidualImage resid;
  resid.image = slot;
  resid.scale = lg.scale;
  if (e == hipSuccess)
    e = hd::vp_poisson(op.xdev, op.mf->nd, op.rho, op.phi, nullptr,
                       op.efield, op.work, op.poisson_tol, op.poisson_max_it,
                       op.stats, s, op.ptab, &resid);
  if (e == hipSuccess && lg.cnn.net.steps > 0)
    {
      e = lg.cnn.net.forward(slot, lg.img_out, 1, s);
      if (e == hipSuccess)
        e = hd::cnn_image_to_dofs(lg.img_out, lg.scale, lg.n0, lg.n1, lg.N,
                                  lg.guess, s);
      if (e == hipSuccess)
        {
          *guess = lg.guess;
          ++lg.guessed;
        }
    }
  if (e != hipSuccess)
    return e;
  // training on the complete pairs (every slot but the head, which this
  // solve fills), overlapped with the CG
  const long long n = lg.solves + 1;
  const int every = n <= lg.dense_solves ? 1 : lg.train_every;
  const bool full = lg.count == lg.capacity;
  const int complete = full ? lg.capacity - 1 : lg.count;
  if (every <= 0 || complete < lg.min_pairs || complete < 1 || n % every != 0)
    return hipSuccess;
  e = hipEventRecord(lg.ev_fork, s);
  if (e == hipSuccess)
    e = hipStreamWaitEvent(lg.train_stream, lg.ev_fork, 0);
  for (int t = 0; t < lg.train_steps && e == hipSuccess; ++t)
    {
      e = hd::cnn_sample(lg.ring_in, lg.ring_tgt, lg.seed, lg.draws, complete,
                         full ? lg.head : -1, lg.train_batch, lg.plane,
                         lg.batch_in, lg.batch_tgt, lg.train_stream);
      lg.draws += (uint64_t)lg.train_batch;
      if (e == hipSuccess)
        e = lg.cnn.net.loss_and_grad(lg.batch_in, lg.batch_tgt,
                                     lg.train_batch, lg.train_stream);
      if (e == hipSuccess)
        e = lg.cnn.net.adam(lg.lr, lg.train_stream);
    }
  if (e == hipSuccess)
    e = hipEventRecord(lg.ev_done, lg.train_stream);
  if (e == hipSuccess)
    lg.pending = true;
  return e;
}

hipError_t learned_after_solve(hd_advection &op, hipStream_t s)
{
  hd_learned_guess &lg = *op.learned;
  // the target image on the training stream too: phi and phi_prev stay as
  // they are until the next solve, which joins that stream first
  hipError_t e = hipEventRecord(lg.ev_fork, s);
  if (e == hipSuccess)
    e = hipStreamWaitEvent(lg.train_stream, lg.ev_fork, 0);
  if (e == hipSuccess)
    e = hd::cnn_dofs_to_image(op.phi, lg.phi_prev, lg.n0, lg.n1, lg.N,
                              lg.ring_tgt + (long long)lg.head * lg.plane,
                              lg.scale_t, lg.train_stream);
  if (e == hipSuccess)
    e = hipEventRecord(lg.ev_done, lg.train_stream);
  if (e != hipSuccess)
    return e;
  lg.pending = true;
  lg.head = (lg.head + 1) % lg.capacity;
  lg.count = std::min(lg.count + 1, lg.capacity);
  ++lg.solves;
  return hipSuccess;
}

void learned_release(hd_learned_guess *lg)
{
  if (!lg)
    return;
  if (lg->train_stream)
    (void)hipStreamSynchronize(lg->train_stream);
  lg->cnn.net.release();
  for (void *p : {(void *)lg->ring_in, (void *)lg->ring_tgt,
                  (void *)lg->img_out, (void *)lg->batch_in,
                  (void *)lg->batch_tgt, (void *)lg->scale,
                  (void *)lg->scale_t, (void *)lg->guess,
                  (void *)lg->phi_prev})
    (void)hipFree(p);
  if (lg->ev_fork)
    (void)hipEventDestroy(lg->ev_fork);
  if (lg->ev_done)
    (void)hipEventDestroy(lg->ev_done);
  if (lg->train_stream)
    (void)hipStreamDestroy(lg->train_stream);
  delete lg;
}
} // namespace hd_abi

namespace
{
bool aligned16(const void *p)
{
  return reinterpret_cast<uintptr_t>(p) % 16 == 0;
}

int check_images(const hd_cnn *net, const void *a, const void *b, int batch,
                 const char *where)
{
  if (!net || !a || !b)
    return fail(HD_ERR_INVALID, std::string(where) + ": null argument");
  if (batch < 1)
    return fail(HD_ERR_INVALID, std::string(where) + ": batch must be >= 1");
  if (!aligned16(a) || !aligned16(b))
    return fail(HD_ERR_INVALID,
                std::string(where) + ": images must be 16-byte aligned");
  return HD_OK;
}
} // namespace

extern "C" {

int hd_cnn_n_params(void) { return hd::CnnNet::n_params(); }

int hd_cnn_create(int height, int width, unsigned long long seed,
                  hd_cnn **out)
{
  if (!out)
    return fail(HD_ERR_INVALID, "hd_cnn_create: null argument");
  *out = nullptr;
  if (!hd::CnnNet::shape_ok(height, width))
    return fail(HD_ERR_INVALID,
                "hd_cnn_create: height and width must be positive multiples "
                "of 16");
  auto *net = new (std::nothrow) hd_cnn();
  if (!net)
    return fail(HD_ERR_NOMEM, "out of host memory");
  const hipError_t e = net->net.init(height, width, seed);
  if (e != hipSuccess)
    {
      net->net.release();
      delete net;
      return hip_fail(e, "hd_cnn_create");
    }
  *out = net;
  return ok();
}

int hd_cnn_destroy(hd_cnn *net)
{
  if (!net)
    return ok();
  net->net.release();
  delete net;
  return ok();
}

int hd_cnn_params(hd_cnn *net, float *host, int device_to_host)
{
  if (!net || !host)
    return fail(HD_ERR_INVALID, "hd_cnn_params: null argument");
  const size_t bytes = hd::CnnNet::n_params() * sizeof(float);
  hd::CnnNet &n = net->net;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    {
      if (device_to_host)
        e = hipMemcpy(host, n.p, bytes, hipMemcpyDeviceToHost);
      else
        {
          // new parameters start a new optimisation
          e = hipMemcpy(n.p, host, bytes, hipMemcpyHostToDevice);
          if (e == hipSuccess)
            e = hipMemset(n.m, 0, bytes);
          if (e == hipSuccess)
            e = hipMemset(n.v, 0, bytes);
          if (e == hipSuccess)
            e = n.pack(nullptr);
          if (e == hipSuccess)
            e = hipDeviceSynchronize();
          n.steps = 0;
        }
    }
  return e == hipSuccess ? ok() : hip_fail(e, "hd_cnn_params");
}

int hd_cnn_forward(hd_cnn *net, const float *in, float *out, int batch,
                   void *stream)
{
  const int rc = check_images(net, in, out, batch, "hd_cnn_forward");
  if (rc != HD_OK)
    return rc;
  const hipError_t e = net->net.forward(in, out, batch, as_stream(stream));
  return e == hipSuccess ? ok() : hip_fail(e, "hd_cnn_forward");
}

int hd_cnn_loss_grad(hd_cnn *net, const float *in, const float *target,
                     int batch, double *loss, float *grad, void *stream)
{
  const int rc = check_images(net, in, target, batch, "hd_cnn_loss_grad");
  if (rc != HD_OK)
    return rc;
  hipStream_t s = as_stream(stream);
  hipError_t e = net->net.loss_and_grad(in, target, batch, s);
  if (e == hipSuccess)
    e = hipStreamSynchronize(s);
  if (e == hipSuccess && loss)
    e = hipMemcpy(loss, net->net.loss_dev, sizeof(double),
                  hipMemcpyDeviceToHost);
  if (e == hipSuccess && grad)
    e = hipMemcpy(grad, net->net.grad,
                  hd::CnnNet::n_params() * sizeof(float),
                  hipMemcpyDeviceToHost);
  return e == hipSuccess ? ok() : hip_fail(e, "hd_cnn_loss_grad");
}

int hd_cnn_train_step(hd_cnn *net, const float *in, const float *target,
                      int batch, float learning_rate, void *stream)
{
  const int rc = check_images(net, in, target, batch, "hd_cnn_train_step");
  if (rc != HD_OK)
    return rc;
  if (!(learning_rate > 0.0f))
    return fail(HD_ERR_INVALID, "hd_cnn_train_step: learning rate <= 0");
  hipStream_t s = as_stream(stream);
  hipError_t e = net->net.loss_and_grad(in, target, batch, s);
  if (e == hipSuccess)
    e = net->net.adam(learning_rate, s);
  return e == hipSuccess ? ok() : hip_fail(e, "hd_cnn_train_step");
}

int hd_cnn_last_loss(hd_cnn *net, double *loss)
{
  if (!net || !loss)
    return fail(HD_ERR_INVALID, "hd_cnn_last_loss: null argument");
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess)
    e = hipMemcpy(loss, net->net.loss_dev, sizeof(double),
                  hipMemcpyDeviceToHost);
  return e == hipSuccess ? ok() : hip_fail(e, "hd_cnn_last_loss");
}

int hd_cnn_dofs_to_image(const double *u, const double *sub, int n0, int n1,
                         int N, float *image, double *scale, void *stream)
{
  if (!u || !image || !scale || n0 < 1 || n1 < 1 || N < 1 ||
      (long long)n0 * n1 * N * N > 0x7fffffffLL)
    return fail(HD_ERR_INVALID, "hd_cnn_dofs_to_image: invalid argument");
  const hipError_t e =
    hd::cnn_dofs_to_image(u, sub, n0, n1, N, image, scale, as_stream(stream));
  return e == hipSuccess ? ok() : hip_fail(e, "hd_cnn_dofs_to_image");
}

int hd_cnn_image_to_dofs(const float *image, const double *scale, int n0,
                         int n1, int N, double *u, void *stream)
{
  if (!u || !image || !scale || n0 < 1 || n1 < 1 || N < 1 ||
      (long long)n0 * n1 * N * N > 0x7fffffffLL)
    return fail(HD_ERR_INVALID, "hd_cnn_image_to_dofs: invalid argument");
  const hipError_t e =
    hd::cnn_image_to_dofs(image, scale, n0, n1, N, u, as_stream(stream));
  return e == hipSuccess ? ok() : hip_fail(e, "hd_cnn_image_to_dofs");
}

int hd_vp_learned_guess_enable(hd_advection *op, unsigned long long seed,
                               int capacity, int train_every, int train_batch,
                               int train_steps, float learning_rate,
                               int min_pairs)
{
  if (!op || !op->vp)
    return fail(HD_ERR_INVALID, "hd_vp_learned_guess_enable: not a "
                                "Vlasov-Poisson operator");
  const hd_matrix_free *mf = op->mf;
  if (mf->dim_x != 2)
    return fail(HD_ERR_UNSUPPORTED,
                "the learned guess works on 2D x-spaces (an image of nodes)");
  const int n0 = mf->mesh.n[0], n1 = mf->mesh.n[1], N = mf->nd;
  if (!op->ptab)
    return fail(HD_ERR_UNSUPPORTED,
                "the learned guess needs the 2D tensor-product Poisson "
                "solver (k <= 3, at most 1024 x cells)");
  if (!hd::CnnNet::shape_ok(n1 * N, n0 * N))
    return fail(HD_ERR_UNSUPPORTED,
                "the learned guess needs (cells x nodes) per x direction to "
                "be a multiple of 16");
  if (capacity < 1 || train_batch < 1 || train_steps < 0 || train_every < 0 ||
      min_pairs < 1 || !(learning_rate > 0.0f))
    return fail(HD_ERR_INVALID, "hd_vp_learned_guess_enable: invalid "
                                "training parameters");
  hd_abi::learned_release(op->learned);
  op->learned = nullptr;
  auto *lg = new (std::nothrow) hd_learned_guess();
  if (!lg)
    return fail(HD_ERR_NOMEM, "out of host memory");
  lg->n0 = n0;
  lg->n1 = n1;
  lg->N = N;
  lg->plane = (long long)n0 * n1 * N * N;
  lg->capacity = capacity;
  lg->seed = seed;
  lg->train_every = train_every;
  lg->train_batch = train_batch;
  lg->train_steps = train_steps;
  lg->lr = learning_rate;
  lg->min_pairs = min_pairs;
  const size_t img = (size_t)lg->plane * sizeof(float);
  hipError_t e = lg->cnn.net.init(n1 * N, n0 * N, seed);
  if (e == hipSuccess)
    e = hipStreamCreateWithFlags(&lg->train_stream, hipStreamNonBlocking);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&lg->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipEventCreateWithFlags(&lg->ev_done, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipMalloc(&lg->ring_in, img * capacity);
  if (e == hipSuccess)
    e = hipMalloc(&lg->ring_tgt, img * capacity);
  if (e == hipSuccess)
    e = hipMalloc(&lg->img_out, img);
  if (e == hipSuccess)
    e = hipMalloc(&lg->batch_in, img * train_batch);
  if (e == hipSuccess)
    e = hipMalloc(&lg->batch_tgt, img * train_batch);
  if (e == hipSuccess)
    e = hipMalloc(&lg->scale, sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&lg->scale_t, sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&lg->guess, (size_t)op->n_x_dofs * sizeof(double));
  if (e == hipSuccess)
    e = hipMalloc(&lg->phi_prev, (size_t)op->n_x_dofs * sizeof(double));
  // the activation buffers for training and inference, up front
  if (e == hipSuccess)
    e = lg->cnn.net.reserve(train_batch);
  if (e != hipSuccess)
    {
      hd_abi::learned_release(lg);
      return hip_fail(e, "hd_vp_learned_guess_enable");
    }
  op->learned = lg;
  op->fdm_on = false; // the two starts exclude each other
  return ok();
}

int hd_vp_learned_guess_schedule(hd_advection *op, long long dense_solves,
                                 int train_every)
{
  if (!op || !op->learned)
    return fail(HD_ERR_INVALID, "hd_vp_learned_guess_schedule: no learned "
                                "guess on this operator");
  if (dense_solves < 0 || train_every < 0)
    return fail(HD_ERR_INVALID, "hd_vp_learned_guess_schedule: negative "
                                "count");
  op->learned->dense_solves = dense_solves;
  op->learned->train_every = train_every;
  return ok();
}

int hd_vp_learned_guess_disable(hd_advection *op)
{
  if (!op || !op->vp)
    return fail(HD_ERR_INVALID, "hd_vp_learned_guess_disable: not a "
                                "Vlasov-Poisson operator");
  const hipError_t e = hipDeviceSynchronize();
  hd_abi::learned_release(op->learned);
  op->learned = nullptr;
  return e == hipSuccess ? ok() : hip_fail(e, "hd_vp_learned_guess_disable");
}

hd_cnn *hd_vp_learned_guess_net(hd_advection *op)
{
  return op && op->learned ? &op->learned->cnn : nullptr;
}

int hd_vp_learned_guess_stats(hd_advection *op, long long *train_steps,
                              long long *guessed_solves, long long *pairs,
                              double *last_loss)
{
  if (!op || !op->learned)
    return fail(HD_ERR_INVALID, "hd_vp_learned_guess_stats: no learned "
                                "guess on this operator");
  const hd_learned_guess &lg = *op->learned;
  if (train_steps)
    *train_steps = lg.cnn.net.steps;
  if (guessed_solves)
    *guessed_solves = lg.guessed;
  if (pairs)
    *pairs = lg.solves;
  if (last_loss)
    {
      *last_loss = 0.0;
      if (lg.cnn.net.steps > 0)
        {
          hipError_t e = hipDeviceSynchronize();
          if (e == hipSuccess)
            e = hipMemcpy(last_loss, lg.cnn.net.loss_dev, sizeof(double),
                          hipMemcpyDeviceToHost);
          if (e != hipSuccess)
            return hip_fail(e, "hd_vp_learned_guess_stats");
        }
    }
  return ok();
}

} // extern "C"
