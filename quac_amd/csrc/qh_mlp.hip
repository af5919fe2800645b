// qh_mlp.hip -- the small fully connected network that predicts a warm start
// for the steady-state solve: forward, backward and Adam, all on the device.
//
// Network: x (in) -> [Linear -> tanh] x (L-1) -> Linear -> y (out), FP32.
// Parameters live in one contiguous FP32 array, layer by layer: W_l (dout x din,
// row-major) then b_l (dout).  Activations are stored row-major per sample
// (batch x width).  Loss = mean over the batch of sum over outputs of
// 0.5 (y - t)^2 weighted by a per-output weight w (the steady-state pipeline
// weights each coefficient by the singular value of its basis vector).
//
// One tiled GEMM kernel serves the three products of a layer:
//   forward   Z = H W^T + b          (batch x dout)
//   backward  dW = dZ^T H            (dout x din)
//             dH = dZ W              (batch x din)
// with op(A), op(B) given by strides; 64 x 64 output tiles per 256-thread
// block, each thread computing a 4 x 4 micro-tile from 64 x 16 LDS tiles.
// The network is tiny next to the solves it serves (a few MFLOP per sample),
// so plain FP32 VALU FMAs are used rather than MFMA tiles.
#include <cmath>
#include <cstring>
#include <vector>

#include "qh_common.h"

struct qh_mlp {
  int nl = 0;                    // number of linear layers
  std::vector<int> dims;         // nl + 1 widths
  std::vector<int64_t> woff, boff;
  int64_t nparam = 0;
  int maxb = 0;                  // batch capacity of the workspaces
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // params, grads, Adam moments
  float *act = nullptr;          // activations: layer l output at act + aoff[l]
  float *dlt = nullptr;          // deltas, same layout
  float *wout = nullptr;         // per-output loss weights
  float *loss = nullptr;         // device scalar(s)
  std::vector<int64_t> aoff;
  int64_t step = 0;
};

namespace {

const int TM = 64, TN = 64, TK = 16;

// C[i,j] = alpha * sum_k A(i,k) B(k,j) + beta * C[i,j] (+ bias[j]) with
// A(i,k) = A[i*sai + k*sak], B(k,j) = B[k*sbk + j*sbj], C row-major (ldc).
// act: 0 none, 1 tanh (applied after the bias)
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, float alpha, const float *__restrict__ A,
                                              int64_t sai, int64_t sak, const float *__restrict__ B, int64_t sbk,
                                              int64_t sbj, float beta, float *__restrict__ C, int64_t ldc,
                                              const float *__restrict__ bias, int act) {
  __shared__ float As[TK][TM + 1];
  __shared__ float Bs[TK][TN + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int i0 = blockIdx.y * TM, j0 = blockIdx.x * TN;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += TK) {
    // 64x16 tile of A and 16x64 tile of B: 1024 elements each, 4 per thread
    for (int e = threadIdx.x; e < TM * TK; e += 256) {
      const int ii = e / TK, kk = e % TK;
      const int gi = i0 + ii, gk = k0 + kk;
      As[kk][ii] = (gi < M && gk < K) ? A[gi * sai + gk * sak] : 0.f;
    }
    for (int e = threadIdx.x; e < TK * TN; e += 256) {
      const int kk = e / TN, jj = e % TN;
      const int gk = k0 + kk, gj = j0 + jj;
      Bs[kk][jj] = (gk < K && gj < N) ? B[gk * sbk + gj * sbj] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TK; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        a[u] = As[kk][ty * 4 + u];
        b[u] = Bs[kk][tx * 4 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int w = 0; w < 4; w++) acc[u][w] = fmaf(a[u], b[w], acc[u][w]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int gi = i0 + ty * 4 + u;
    if (gi >= M) continue;
#pragma unroll
    for (int w = 0; w < 4; w++) {
      const int gj = j0 + tx * 4 + w;
      if (gj >= N) continue;
      float c = alpha * acc[u][w];
      if (beta != 0.f) c += beta * C[gi * ldc + gj];
      if (bias) c += bias[gj];
      if (act == 1) c = tanhf(c);
      C[gi * ldc + gj] = c;
    }
  }
}

void gemm(hipStream_t s, int M, int N, int K, float alpha, const float *A, int64_t sai, int64_t sak, const float *B,
          int64_t sbk, int64_t sbj, float beta, float *C, int64_t ldc, const float *bias, int act) {
  dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
  hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, M, N, K, alpha, A, sai, sak, B, sbk, sbj, beta, C, ldc, bias,
                     act);
}

// output delta of the weighted squared loss, and the per-sample loss partials:
// d = w (y - t) / B ; part[b] = sum_j 0.5 w_j (y - t)^2
__global__ void k_out_delta(const float *y, const float *t, const float *w, int B, int n, float *d, float *part) {
  const int b = blockIdx.x;
  __shared__ float red[256];
  float acc = 0.f;
  for (int j = threadIdx.x; j < n; j += 256) {
    const float e = y[b * n + j] - t[b * n + j];
    d[b * n + j] = w[j] * e / (float)B;
    acc += 0.5f * w[j] * e * e;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[b] = red[0];
}

// loss = sum(part) / B
__global__ void k_loss_sum(const float *part, int B, float *loss) {
  __shared__ float red[256];
  float acc = 0.f;
  for (int b = threadIdx.x; b < B; b += 256) acc += part[b];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = red[0] / (float)B;
}

// delta through tanh: d *= 1 - h^2 (h = the layer's tanh output)
__global__ void k_tanh_back(float *d, const float *h, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] *= 1.f - h[i] * h[i];
}

// db_j = sum_b d[b, j]
__global__ void k_colsum(const float *d, int B, int n, float *db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float acc = 0.f;
  for (int b = 0; b < B; b++) acc += d[b * n + j];
  db[j] = acc;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

__global__ void k_zero_f(float *a, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = 0.f;
}

unsigned g1(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  return (unsigned)(b < 1 ? 1 : b);
}

// forward pass of B samples (x: B x in, device); activations into net->act
int forward(qh_mlp *net, const float *x, int B) {
  hipStream_t s = qh_stream();
  const float *h = x;
  for (int l = 0; l < net->nl; l++) {
    const int din = net->dims[l], dout = net->dims[l + 1];
    float *z = net->act + net->aoff[l + 1];
    // Z = H W^T + b:  A = H (B x din, sai = din, sak = 1), B(k, j) = W[j, k] (sbk = 1, sbj = din)
    gemm(s, B, dout, din, 1.f, h, din, 1, net->p + net->woff[l], 1, din, 0.f, z, dout, net->p + net->boff[l],
         l + 1 < net->nl ? 1 : 0);
    h = z;
  }
  QH_HIP(hipGetLastError());
  return QH_OK;
}

// gradients of the weighted loss into net->g (after forward on the same x)
int backward(qh_mlp *net, const float *x, const float *t, int B) {
  hipStream_t s = qh_stream();
  const int L = net->nl;
  const int nout = net->dims[L];
  float *part = net->loss + 1;
  hipLaunchKernelGGL(k_out_delta, dim3(B), dim3(256), 0, s, net->act + net->aoff[L], t, net->wout, B, nout,
                     net->dlt + net->aoff[L], part);
  hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(256), 0, s, part, B, net->loss);
  for (int l = L - 1; l >= 0; l--) {
    const int din = net->dims[l], dout = net->dims[l + 1];
    const float *hin = l == 0 ? x : net->act + net->aoff[l];
    float *d = net->dlt + net->aoff[l + 1];
    // dW = dZ^T H: C (dout x din) ; A(i=o, k=b) = d[b*dout + o] (sai = 1, sak = dout); B(k=b, j) = hin[b*din + j]
    gemm(s, dout, din, B, 1.f, d, 1, dout, hin, din, 1, 0.f, net->g + net->woff[l], din, nullptr, 0);
    hipLaunchKernelGGL(k_colsum, dim3((dout + 255) / 256), dim3(256), 0, s, d, B, dout, net->g + net->boff[l]);
    if (l > 0) {
      float *dh = net->dlt + net->aoff[l];
      // dH = dZ W: A = d (B x dout), B(k=o, j) = W[o*din + j]
      gemm(s, B, din, dout, 1.f, d, dout, 1, net->p + net->woff[l], din, 1, 0.f, dh, din, nullptr, 0);
      hipLaunchKernelGGL(k_tanh_back, dim3(g1((int64_t)B * din)), dim3(256), 0, s, dh, net->act + net->aoff[l],
                         (int64_t)B * din);
    }
  }
  QH_HIP(hipGetLastError());
  return QH_OK;
}

}  // namespace

extern "C" int qh_mlp_destroy(qh_mlp *net) {
  if (!net) return QH_OK;
  qh_free_ptr(net->p);
  qh_free_ptr(net->g);
  qh_free_ptr(net->m);
  qh_free_ptr(net->v);
  qh_free_ptr(net->act);
  qh_free_ptr(net->dlt);
  qh_free_ptr(net->wout);
  qh_free_ptr(net->loss);
  delete net;
  return QH_OK;
}

extern "C" int qh_mlp_create(int nlayers, const int *dims, int max_batch, qh_mlp **out) {
  int rc = qh_ensure_init();
  if (rc) return rc;
  QH_REQUIRE(out && dims && nlayers >= 1 && nlayers <= 16 && max_batch >= 1 && max_batch <= (1 << 20), QH_EINVAL,
             "qh_mlp_create: 1..16 layers and a batch capacity in [1, 2^20] required");
  for (int l = 0; l <= nlayers; l++)
    QH_REQUIRE(dims[l] >= 1 && dims[l] <= 65536, QH_EINVAL, "qh_mlp_create: widths in [1, 65536] required");
  qh_mlp *net = new qh_mlp();
  net->nl = nlayers;
  net->dims.assign(dims, dims + nlayers + 1);
  net->maxb = max_batch;
  int64_t off = 0;
  for (int l = 0; l < nlayers; l++) {
    net->woff.push_back(off);
    off += (int64_t)dims[l] * dims[l + 1];
    net->boff.push_back(off);
    off += dims[l + 1];
  }
  net->nparam = off;
  int64_t aoff = 0;
  for (int l = 0; l <= nlayers; l++) {
    net->aoff.push_back(aoff);
    aoff += (int64_t)max_batch * dims[l];
  }
  rc = qh_alloc_t(&net->p, off);
  if (!rc) rc = qh_alloc_t(&net->g, off);
  if (!rc) rc = qh_alloc_t(&net->m, off);
  if (!rc) rc = qh_alloc_t(&net->v, off);
  if (!rc) rc = qh_alloc_t(&net->act, aoff);
  if (!rc) rc = qh_alloc_t(&net->dlt, aoff);
  if (!rc) rc = qh_alloc_t(&net->wout, dims[nlayers]);
  if (!rc) rc = qh_alloc_t(&net->loss, 1 + max_batch);
  if (rc) {
    qh_mlp_destroy(net);
    return rc;
  }
  hipStream_t s = qh_stream();
  for (float *a : {net->p, net->g, net->m, net->v})
    hipLaunchKernelGGL(k_zero_f, dim3(g1(off)), dim3(256), 0, s, a, off);
  std::vector<float> ones(dims[nlayers], 1.f);
  QH_HIP(hipMemcpyAsync(net->wout, ones.data(), sizeof(float) * ones.size(), hipMemcpyHostToDevice, s));
  QH_HIP(hipStreamSynchronize(s));
  *out = net;
  return QH_OK;
}

extern "C" int qh_mlp_num_params(const qh_mlp *net, int64_t *n) {
  QH_REQUIRE(net && n, QH_EINVAL, "null argument");
  *n = net->nparam;
  return QH_OK;
}

// host <-> device parameter copies (layout: W_0, b_0, W_1, b_1, ...)
extern "C" int qh_mlp_set_params(qh_mlp *net, const float *host) {
  QH_REQUIRE(net && host, QH_EINVAL, "null argument");
  hipStream_t s = qh_stream();
  QH_HIP(hipMemcpyAsync(net->p, host, sizeof(float) * net->nparam, hipMemcpyHostToDevice, s));
  // a new starting point restarts the optimiser
  hipLaunchKernelGGL(k_zero_f, dim3(g1(net->nparam)), dim3(256), 0, s, net->m, net->nparam);
  hipLaunchKernelGGL(k_zero_f, dim3(g1(net->nparam)), dim3(256), 0, s, net->v, net->nparam);
  net->step = 0;
  QH_HIP(hipStreamSynchronize(s));
  return QH_OK;
}

extern "C" int qh_mlp_get_params(const qh_mlp *net, float *host) {
  QH_REQUIRE(net && host, QH_EINVAL, "null argument");
  QH_HIP(hipMemcpyAsync(host, net->p, sizeof(float) * net->nparam, hipMemcpyDeviceToHost, qh_stream()));
  QH_HIP(hipStreamSynchronize(qh_stream()));
  return QH_OK;
}

extern "C" int qh_mlp_set_output_weights(qh_mlp *net, const float *host) {
  QH_REQUIRE(net && host, QH_EINVAL, "null argument");
  QH_HIP(hipMemcpyAsync(net->wout, host, sizeof(float) * net->dims[net->nl], hipMemcpyHostToDevice, qh_stream()));
  QH_HIP(hipStreamSynchronize(qh_stream()));
  return QH_OK;
}

// y (B x out) = net(x), x: B x in; both device pointers
extern "C" int qh_mlp_forward(qh_mlp *net, const float *x, int B, float *y) {
  QH_REQUIRE(net && x && y && B >= 1 && B <= net->maxb, QH_EINVAL, "qh_mlp_forward: batch in [1, capacity]");
  int rc = forward(net, x, B);
  if (rc) return rc;
  const int nout = net->dims[net->nl];
  QH_HIP(hipMemcpyAsync(y, net->act + net->aoff[net->nl], sizeof(float) * B * nout, hipMemcpyDeviceToDevice,
                        qh_stream()));
  return QH_OK;
}

// loss and gradients (host copies) at the current parameters for a batch (device x, t)
extern "C" int qh_mlp_loss_grad(qh_mlp *net, const float *x, const float *t, int B, float *loss, float *grad_host) {
  QH_REQUIRE(net && x && t && B >= 1 && B <= net->maxb, QH_EINVAL, "qh_mlp_loss_grad: batch in [1, capacity]");
  int rc = forward(net, x, B);
  if (!rc) rc = backward(net, x, t, B);
  if (rc) return rc;
  hipStream_t s = qh_stream();
  if (loss) QH_HIP(hipMemcpyAsync(loss, net->loss, sizeof(float), hipMemcpyDeviceToHost, s));
  if (grad_host) QH_HIP(hipMemcpyAsync(grad_host, net->g, sizeof(float) * net->nparam, hipMemcpyDeviceToHost, s));
  QH_HIP(hipStreamSynchronize(s));
  return QH_OK;
}

// one Adam step on a batch; loss (host, optional) is the batch loss before the step
extern "C" int qh_mlp_train_step(qh_mlp *net, const float *x, const float *t, int B, float lr, float beta1,
                                 float beta2, float eps, float *loss) {
  QH_REQUIRE(net && x && t && B >= 1 && B <= net->maxb, QH_EINVAL, "qh_mlp_train_step: batch in [1, capacity]");
  int rc = forward(net, x, B);
  if (!rc) rc = backward(net, x, t, B);
  if (rc) return rc;
  net->step++;
  const float c1 = 1.f - powf(beta1, (float)net->step), c2 = 1.f - powf(beta2, (float)net->step);
  hipStream_t s = qh_stream();
  hipLaunchKernelGGL(k_adam, dim3(g1(net->nparam)), dim3(256), 0, s, net->p, net->g, net->m, net->v, net->nparam, lr,
                     beta1, beta2, eps, c1, c2);
  QH_HIP(hipGetLastError());
  if (loss) {
    QH_HIP(hipMemcpyAsync(loss, net->loss, sizeof(float), hipMemcpyDeviceToHost, s));
    QH_HIP(hipStreamSynchronize(s));
  }
  return QH_OK;
}
