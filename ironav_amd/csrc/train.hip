// train.hip — FP32 training kernels of the motion-proposal CNN (offline tooling; libironav_train.so).
//
// The network is the one proposal.hip runs in BF16 on MFMA: four 3x3 convolutions on a 9x9 grid
// with zero padding, ReLU after the first three. Tensors: activations [batch][81][channels],
// weights [cout][cin][3][3] (the layout ia_load_proposal_weights takes), biases [cout].
// Loss per sample and reference with a target cell: softmax cross-entropy of the reference's score
// map plus cost_w * (predicted log2 SAD at the target cell - target)^2; averaged over the batch.
// Every entry point takes device pointers and the HIP stream to enqueue on.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int G = 9, CELLS = 81, OUT = 4, NREF = 3;

__device__ __forceinline__ int cell_at(int p, int t) {
  const int y = p / G + t / 3 - 1, x = p % G + t % 3 - 1;
  return (y < 0 || y >= G || x < 0 || x >= G) ? -1 : y * G + x;
}

// out[n][p][co] = b[co] + sum_{ci,t} in[n][p+t][ci] * w[co][ci][t]  (optionally ReLU)
__global__ void conv_fwd(const float *in, const float *w, const float *b, float *out, int B, int cin, int cout, int relu) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)B * CELLS * cout) return;
  const int co = (int)(i % cout), p = (int)((i / cout) % CELLS);
  const long n = i / ((long)cout * CELLS);
  float s = b[co];
  for (int t = 0; t < 9; t++) {
    const int q = cell_at(p, t);
    if (q < 0) continue;
    const float *x = in + (n * CELLS + q) * cin;
    const float *wt = w + (long)co * cin * 9 + t;
    for (int ci = 0; ci < cin; ci++) s += x[ci] * wt[ci * 9];
  }
  out[i] = relu ? fmaxf(s, 0.f) : s;
}

// din[n][q][ci] = sum_{co,t} dout[n][q - t][co] * w[co][ci][t]
__global__ void conv_bwd_data(const float *dout, const float *w, float *din, int B, int cin, int cout) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)B * CELLS * cin) return;
  const int ci = (int)(i % cin), q = (int)((i / cin) % CELLS);
  const long n = i / ((long)cin * CELLS);
  float s = 0.f;
  for (int t = 0; t < 9; t++) {
    // output cell p whose tap t reads q: p + off(t) = q  <=>  p = q - off(t) = cell_at(q, 8 - t)
    const int p = cell_at(q, 8 - t);
    if (p < 0) continue;
    const float *d = dout + (n * CELLS + p) * cout;
    for (int co = 0; co < cout; co++) s += d[co] * w[((long)co * cin + ci) * 9 + t];
  }
  din[i] = s;
}

// dw[co][ci][t] += sum_{n,p} in[n][p+t][ci] * dout[n][p][co]; db[co] += sum dout (chunked over n)
__global__ void conv_bwd_w(const float *in, const float *dout, float *dw, float *db, int B, int cin, int cout, int chunk) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nw = (long)cout * cin * 9;
  const int nchunks = (B + chunk - 1) / chunk;
  if (i >= (nw + cout) * nchunks) return;
  const int c = (int)(i % nchunks);
  const long k = i / nchunks;
  const int n0 = c * chunk, n1 = min(B, n0 + chunk);
  float s = 0.f;
  if (k < nw) {
    const int t = (int)(k % 9), ci = (int)((k / 9) % cin), co = (int)(k / (9L * cin));
    for (int n = n0; n < n1; n++)
      for (int p = 0; p < CELLS; p++) {
        const int q = cell_at(p, t);
        if (q < 0) continue;
        s += in[((long)n * CELLS + q) * cin + ci] * dout[((long)n * CELLS + p) * cout + co];
      }
    atomicAdd(&dw[k], s);
  } else {
    const int co = (int)(k - nw);
    for (int n = n0; n < n1; n++)
      for (int p = 0; p < CELLS; p++) s += dout[((long)n * CELLS + p) * cout + co];
    atomicAdd(&db[co], s);
  }
}

// d *= (act > 0)
__global__ void relu_bwd(float *d, const float *act, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && act[i] <= 0.f) d[i] = 0.f;
}

// one thread per (sample, reference): loss and its gradient into dout[n][p][OUT]
__global__ void loss_grad(const float *out, const int *tcell, const float *tcost, float *dout, float *loss, int B,
                          float cost_w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * NREF) return;
  const int n = i / NREF, r = i % NREF, tc = tcell[i];
  if (tc < 0) return;
  const float *o = out + (long)n * CELLS * OUT;
  float *d = dout + (long)n * CELLS * OUT;
  float mx = -1e30f;
  for (int p = 0; p < CELLS; p++) mx = fmaxf(mx, o[p * OUT + r]);
  float z = 0.f;
  for (int p = 0; p < CELLS; p++) z += expf(o[p * OUT + r] - mx);
  const float inv = 1.f / (z * B);
  for (int p = 0; p < CELLS; p++) d[p * OUT + r] = expf(o[p * OUT + r] - mx) * inv;
  d[tc * OUT + r] -= 1.f / B;
  const float e = o[tc * OUT + 3] - tcost[i];
  atomicAdd(&d[tc * OUT + 3], 2.f * cost_w * e / B);
  atomicAdd(loss, (mx + logf(z) - o[tc * OUT + r] + cost_w * e * e) / B);
}

__global__ void adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                     float c1, float c2, float wd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * p[i];
  m[i] = b1 * m[i] + (1.f - b1) * gi;
  v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
  p[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
}

inline dim3 blocks(long n, int t = 256) { return dim3((unsigned)((n + t - 1) / t)); }

}  // namespace

extern "C" {

int it_conv_fwd(const float *in, const float *w, const float *b, float *out, int B, int cin, int cout, int relu,
                void *stream) {
  hipLaunchKernelGGL(conv_fwd, blocks((long)B * CELLS * cout), dim3(256), 0, (hipStream_t)stream, in, w, b, out, B, cin,
                     cout, relu);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int it_conv_bwd_data(const float *dout, const float *w, float *din, int B, int cin, int cout, void *stream) {
  hipLaunchKernelGGL(conv_bwd_data, blocks((long)B * CELLS * cin), dim3(256), 0, (hipStream_t)stream, dout, w, din, B,
                     cin, cout);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int it_conv_bwd_w(const float *in, const float *dout, float *dw, float *db, int B, int cin, int cout, void *stream) {
  const int chunk = 16, nchunks = (B + chunk - 1) / chunk;
  hipLaunchKernelGGL(conv_bwd_w, blocks(((long)cout * cin * 9 + cout) * nchunks), dim3(256), 0, (hipStream_t)stream,
                     in, dout, dw, db, B, cin, cout, chunk);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int it_relu_bwd(float *d, const float *act, long n, void *stream) {
  hipLaunchKernelGGL(relu_bwd, blocks(n), dim3(256), 0, (hipStream_t)stream, d, act, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int it_loss_grad(const float *out, const int *tcell, const float *tcost, float *dout, float *loss, int B, float cost_w,
                 void *stream) {
  hipLaunchKernelGGL(loss_grad, blocks((long)B * NREF), dim3(256), 0, (hipStream_t)stream, out, tcell, tcost, dout,
                     loss, B, cost_w);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int it_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps, int step,
            float wd, void *stream) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam, blocks(n), dim3(256), 0, (hipStream_t)stream, p, g, m, v, n, lr, b1, b2, eps, c1, c2, wd);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
