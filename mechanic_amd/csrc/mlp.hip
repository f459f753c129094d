// Learned warm start: inference over host-supplied feature rows and GPU
// training of the network in mlp_net.h.
//
// Training runs in FP32 with this file's forward, backward and Adam kernels:
// each block takes kChunk rows, keeps their activations in LDS and writes its
// partial gradient; the Adam kernel sums the partials in block order (so a
// step is deterministic) and updates the FP32 master weights.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/mechanic_hip.h"
#include "internal.h"
#include "mlp_net.h"

namespace mh {

__global__ __launch_bounds__(256) void k_mlp_predict(const float* net, const float* feat, uint32_t rows, float* out) {
  MlpRegs m;
  const uint32_t lane = threadIdx.x & 63u;
  mlp_load(net, lane, m);
  const uint32_t groups = (rows + 15) / 16;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < groups; g += waves) {
    const uint32_t c = g * 16 + lane % 16;
    float x[4] = {0.f, 0.f, 0.f, 0.f};
    if (c < rows) {
      const float4 v = *reinterpret_cast<const float4*>(feat + (size_t)c * kFeatures + 4 * (lane / 16));
      x[0] = v.x;
      x[1] = v.y;
      x[2] = v.z;
      x[3] = v.w;
    }
    float y[3];
    mlp_group(m, x, y);
    if (lane < 16 && c < rows) {
      out[3 * (size_t)c] = y[0];
      out[3 * (size_t)c + 1] = y[1];
      out[3 * (size_t)c + 2] = y[2];
    }
  }
}

// ---------------------------------------------------------------------------
// Training: FP32 forward and backward of a chunk of kChunk contacts per block.
// Loss = mean over rows of the squared error of the normalised outputs.
// ---------------------------------------------------------------------------
constexpr int kChunk = 32;

// `order` maps batch rows to dataset rows; null takes the rows in dataset order.
__global__ __launch_bounds__(256) void k_mlp_grad(const float* net, const float* feat, const float* targ,
                                                  const uint32_t* order, uint32_t first, uint32_t rows,
                                                  float inv_batch, float* partial) {
  __shared__ float X[kChunk][kIn];
  __shared__ float Z1[kChunk][kH], H1[kChunk][kH], Z2[kChunk][kH], H2[kChunk][kH];
  __shared__ float DY[kChunk][4];
  __shared__ float LOSS[kChunk];
  const int t = threadIdx.x;
  const uint32_t base = first + blockIdx.x * kChunk;
  // inputs
  for (int e = t; e < kChunk * kIn; e += blockDim.x) {
    const int c = e / kIn, k = e % kIn;
    const uint32_t row = base + c;
    float v = 0.0f;
    if (row < first + rows) {
      const uint32_t src = order ? order[row] : row;
      v = (feat[(size_t)src * kFeatures + k] - net[kOffInMean + k]) * net[kOffInScale + k];
    }
    X[c][k] = v;
  }
  __syncthreads();
  for (int e = t; e < kChunk * kH; e += blockDim.x) {
    const int c = e / kH, h = e % kH;
    float z = net[kOffB1 + h];
    for (int k = 0; k < kIn; ++k) z += X[c][k] * net[kOffW1 + k * kH + h];
    Z1[c][h] = z;
    H1[c][h] = gelu_exact(z);
  }
  __syncthreads();
  for (int e = t; e < kChunk * kH; e += blockDim.x) {
    const int c = e / kH, h = e % kH;
    float z = net[kOffB2 + h];
    for (int k = 0; k < kH; ++k) z += H1[c][k] * net[kOffW2 + k * kH + h];
    Z2[c][h] = z;
    H2[c][h] = gelu_exact(z);
  }
  __syncthreads();
  if (t < kChunk) {
    const uint32_t row = base + t;
    float loss = 0.0f;
    for (int o = 0; o < kOut; ++o) {
      float y = net[kOffB3 + o];
      for (int k = 0; k < kH; ++k) y += H2[t][k] * net[kOffW3 + k * kOut + o];
      float dy = 0.0f;
      if (row < first + rows) {
        const uint32_t src = order ? order[row] : row;
        const float target = (targ[3 * (size_t)src + o] - net[kOffOutMean + o]) / net[kOffOutScale + o];
        const float err = y - target;
        loss += err * err;
        dy = 2.0f * err * inv_batch;
      }
      DY[t][o] = dy;
    }
    LOSS[t] = loss * inv_batch;
  }
  __syncthreads();
  // back through layer 3 and the second GELU: dZ2 reuses Z2
  for (int e = t; e < kChunk * kH; e += blockDim.x) {
    const int c = e / kH, k = e % kH;
    float g = 0.0f;
    for (int o = 0; o < kOut; ++o) g += DY[c][o] * net[kOffW3 + k * kOut + o];
    Z2[c][k] = g * gelu_grad(Z2[c][k]);
  }
  __syncthreads();
  // back through layer 2 and the first GELU: dZ1 reuses Z1
  for (int e = t; e < kChunk * kH; e += blockDim.x) {
    const int c = e / kH, k = e % kH;
    float g = 0.0f;
    for (int h = 0; h < kH; ++h) g += Z2[c][h] * net[kOffW2 + k * kH + h];
    Z1[c][k] = g * gelu_grad(Z1[c][k]);
  }
  __syncthreads();
  float* out = partial + (size_t)blockIdx.x * (kParams + 1);
  for (int p = t; p < kParams; p += blockDim.x) {
    float g = 0.0f;
    if (p < kOffB1) {
      const int k = p / kH, h = p % kH;
      for (int c = 0; c < kChunk; ++c) g += X[c][k] * Z1[c][h];
    } else if (p < kOffW2) {
      const int h = p - kOffB1;
      for (int c = 0; c < kChunk; ++c) g += Z1[c][h];
    } else if (p < kOffB2) {
      const int k = (p - kOffW2) / kH, h = (p - kOffW2) % kH;
      for (int c = 0; c < kChunk; ++c) g += H1[c][k] * Z2[c][h];
    } else if (p < kOffW3) {
      const int h = p - kOffB2;
      for (int c = 0; c < kChunk; ++c) g += Z2[c][h];
    } else if (p < kOffB3) {
      const int k = (p - kOffW3) / kOut, o = (p - kOffW3) % kOut;
      for (int c = 0; c < kChunk; ++c) g += H2[c][k] * DY[c][o];
    } else {
      const int o = p - kOffB3;
      for (int c = 0; c < kChunk; ++c) g += DY[c][o];
    }
    out[p] = g;
  }
  if (t == 0) {
    float l = 0.0f;
    for (int c = 0; c < kChunk; ++c) l += LOSS[c];
    out[kParams] = l;
  }
}

// Sums the block partials in order; optionally applies one Adam step.
__global__ __launch_bounds__(256) void k_mlp_adam(float* net, float* m1, float* m2, const float* partial,
                                                  uint32_t blocks, float lr, float beta1, float beta2, float eps,
                                                  float correction1, float correction2, int apply, float* grad_out,
                                                  float* loss_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kParams) return;
  float g = 0.0f;
  for (uint32_t b = 0; b < blocks; ++b) g += partial[(size_t)b * (kParams + 1) + p];
  if (p == kParams) {
    if (loss_out) *loss_out += g;  // one thread, in stream order: deterministic
    return;
  }
  if (grad_out) grad_out[p] = g;
  if (!apply) return;
  const float a = beta1 * m1[p] + (1.0f - beta1) * g;
  const float v = beta2 * m2[p] + (1.0f - beta2) * g * g;
  m1[p] = a;
  m2[p] = v;
  net[p] -= lr * (a / correction1) / (sqrtf(v / correction2) + eps);
}

// Running sums over windows of block partials: sums[p] (kParams gradient
// entries, then the loss) continues from its previous value -- zero for the
// first window -- and adds the window's partials in block order. Window after
// window this is the same left-to-right sum k_mlp_adam forms over all blocks
// at once, so the result has the same bits.
__global__ __launch_bounds__(256) void k_mlp_fold(const float* partial, uint32_t blocks, int first_window,
                                                  float* grad, float* loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kParams) return;
  float* out = p == kParams ? loss : grad + p;
  float g = first_window ? 0.0f : *out;
  for (uint32_t b = 0; b < blocks; ++b) g += partial[(size_t)b * (kParams + 1) + p];
  *out = g;
}

// Block partials of the allocation-free gradient (mh_mlp_gradient_noalloc):
// one window of kFoldBlocks blocks of kChunk rows (4,096 rows, 2.8 MB). Every
// device the library runs on holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns.
constexpr uint32_t kFoldBlocks = 128;
constexpr size_t kMaxNoallocRows = (size_t)1 << 28;  // the dataset cap of mh_mlp_set_dataset
static __device__ __align__(256) float scratch[kFoldBlocks * (kParams + 1)];

}  // namespace mh

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
using namespace mh;

struct mh_mlp {
  float* net = nullptr;  // kNetFloats
  float* m1 = nullptr;
  float* m2 = nullptr;
  float* partial = nullptr;
  uint32_t partial_blocks = 0;
  float* feat = nullptr;
  float* targ = nullptr;
  uint32_t* order = nullptr;
  float* scratch = nullptr;
  size_t rows = 0;
  uint64_t step = 0;
  uint64_t rng = 0;
  hipStream_t stream = nullptr;
};

namespace {

uint64_t splitmix(uint64_t& s) {
  s += 0x9e3779b97f4a7c15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double uniform(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

#define MLP_TRY(expr)                        \
  do {                                       \
    if ((expr) != hipSuccess) return MH_ERR_DEVICE; \
  } while (0)

int ensure_partials(mh_mlp* m, uint32_t blocks) {
  if (blocks <= m->partial_blocks) return MH_OK;
  if (m->partial) (void)hipFree(m->partial);
  m->partial = nullptr;
  MLP_TRY(hipMalloc(&m->partial, (size_t)blocks * (kParams + 1) * sizeof(float)));
  m->partial_blocks = blocks;
  return MH_OK;
}

}  // namespace

namespace mh {
const float* mlp_device_net(const mh_mlp* m) { return m ? m->net : nullptr; }
}  // namespace mh

extern "C" {

int mh_mlp_create(uint64_t seed, mh_mlp** out) {
  if (!out) return MH_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) return MH_ERR_NO_DEVICE;
  mh_mlp* m = new mh_mlp();
  m->rng = seed;
  std::vector<float> host(kNetFloats, 0.0f);
  // Uniform fan-in initialisation; biases zero; identity normalisation.
  auto fill = [&](int off, int fan_in, int count) {
    const double bound = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < count; ++i) host[off + i] = (float)((2.0 * uniform(m->rng) - 1.0) * bound);
  };
  fill(kOffW1, kIn, kIn * kH);
  fill(kOffW2, kH, kH * kH);
  fill(kOffW3, kH, kH * kOut);
  for (int i = 0; i < kIn; ++i) host[kOffInScale + i] = 1.0f;
  for (int i = 0; i < kOut; ++i) host[kOffOutScale + i] = 1.0f;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&m->net, kNetFloats * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->m1, kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->m2, kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->scratch, (kParams + 1) * sizeof(float)) != hipSuccess ||
      hipMemcpy(m->net, host.data(), kNetFloats * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->m1, 0, kParams * sizeof(float)) != hipSuccess ||
      hipMemset(m->m2, 0, kParams * sizeof(float)) != hipSuccess) {
    mh_mlp_destroy(m);
    return MH_ERR_DEVICE;
  }
  *out = m;
  return MH_OK;
}

void mh_mlp_destroy(mh_mlp* m) {
  if (!m) return;
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : {(void*)m->net, (void*)m->m1, (void*)m->m2, (void*)m->partial, (void*)m->feat, (void*)m->targ,
                  (void*)m->order, (void*)m->scratch})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

uint32_t mh_mlp_param_count(void) { return (uint32_t)kParams; }
uint32_t mh_mlp_net_floats(void) { return (uint32_t)kNetFloats; }

int mh_mlp_get_net(mh_mlp* m, float* out, size_t count) {
  if (!m || !out || count < (size_t)kNetFloats) return MH_ERR_INVALID_ARGUMENT;
  MLP_TRY(hipStreamSynchronize(m->stream));
  MLP_TRY(hipMemcpy(out, m->net, kNetFloats * sizeof(float), hipMemcpyDeviceToHost));
  return MH_OK;
}

int mh_mlp_set_net(mh_mlp* m, const float* net, size_t count) {
  if (!m || !net || count < (size_t)kNetFloats) return MH_ERR_INVALID_ARGUMENT;
  MLP_TRY(hipStreamSynchronize(m->stream));
  MLP_TRY(hipMemcpy(m->net, net, kNetFloats * sizeof(float), hipMemcpyHostToDevice));
  return MH_OK;
}

// Copies a dataset of `rows` contacts (kFeatures inputs, 3 targets each) to the
// device and sets the input/output normalisation to its mean and 1/std (output:
// std) per column.
int mh_mlp_set_dataset(mh_mlp* m, const float* features, const float* targets, size_t rows) {
  if (!m || !features || !targets || rows == 0 || rows > (1u << 28)) return MH_ERR_INVALID_ARGUMENT;
  MLP_TRY(hipStreamSynchronize(m->stream));
  for (void** p : {(void**)&m->feat, (void**)&m->targ, (void**)&m->order}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  MLP_TRY(hipMalloc(&m->feat, rows * kFeatures * sizeof(float)));
  MLP_TRY(hipMalloc(&m->targ, rows * kOut * sizeof(float)));
  MLP_TRY(hipMalloc(&m->order, rows * sizeof(uint32_t)));
  MLP_TRY(hipMemcpy(m->feat, features, rows * kFeatures * sizeof(float), hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(m->targ, targets, rows * kOut * sizeof(float), hipMemcpyHostToDevice));
  m->rows = rows;
  std::vector<float> net(kNetFloats);
  MLP_TRY(hipMemcpy(net.data(), m->net, kNetFloats * sizeof(float), hipMemcpyDeviceToHost));
  for (int k = 0; k < kIn; ++k) {
    double s = 0.0, s2 = 0.0;
    for (size_t r = 0; r < rows; ++r) {
      const double v = features[r * kFeatures + k];
      s += v;
      s2 += v * v;
    }
    const double mean = s / (double)rows, var = s2 / (double)rows - mean * mean;
    const double sd = var > 1e-12 ? std::sqrt(var) : 1.0;
    net[kOffInMean + k] = (float)mean;
    net[kOffInScale + k] = (float)(1.0 / sd);
  }
  for (int o = 0; o < kOut; ++o) {
    double s = 0.0, s2 = 0.0;
    for (size_t r = 0; r < rows; ++r) {
      const double v = targets[r * kOut + o];
      s += v;
      s2 += v * v;
    }
    const double mean = s / (double)rows, var = s2 / (double)rows - mean * mean;
    net[kOffOutMean + o] = (float)mean;
    net[kOffOutScale + o] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
  }
  MLP_TRY(hipMemcpy(m->net, net.data(), kNetFloats * sizeof(float), hipMemcpyHostToDevice));
  return MH_OK;
}

// Mini-batch Adam over the dataset for `epochs` epochs (shuffled each epoch with
// the handle's seeded generator). losses[e] receives epoch e's mean batch loss.
int mh_mlp_train(mh_mlp* m, uint32_t epochs, uint32_t batch, float lr, float* losses) {
  if (!m || m->rows == 0 || batch == 0) return MH_ERR_INVALID_ARGUMENT;
  const uint32_t rows = (uint32_t)m->rows;
  const uint32_t blocks_max = (batch + kChunk - 1) / kChunk;
  if (ensure_partials(m, blocks_max) != MH_OK) return MH_ERR_DEVICE;
  std::vector<uint32_t> order(rows);
  float* loss_dev = m->scratch + kParams;
  for (uint32_t e = 0; e < epochs; ++e) {
    for (uint32_t i = 0; i < rows; ++i) order[i] = i;
    for (uint32_t i = rows - 1; i > 0; --i) {
      const uint32_t j = (uint32_t)(splitmix(m->rng) % (uint64_t)(i + 1));
      std::swap(order[i], order[j]);
    }
    MLP_TRY(hipMemcpyAsync(m->order, order.data(), rows * sizeof(uint32_t), hipMemcpyHostToDevice, m->stream));
    MLP_TRY(hipMemsetAsync(loss_dev, 0, sizeof(float), m->stream));
    uint32_t steps = 0;
    for (uint32_t first = 0; first < rows; first += batch) {
      const uint32_t n = rows - first < batch ? rows - first : batch;
      const uint32_t blocks = (n + kChunk - 1) / kChunk;
      k_mlp_grad<<<blocks, 256, 0, m->stream>>>(m->net, m->feat, m->targ, m->order, first, n, 1.0f / (float)n,
                                               m->partial);
      ++m->step;
      const float c1 = 1.0f - std::pow(0.9f, (float)m->step), c2 = 1.0f - std::pow(0.999f, (float)m->step);
      k_mlp_adam<<<(kParams + 1 + 255) / 256, 256, 0, m->stream>>>(m->net, m->m1, m->m2, m->partial, blocks, lr,
                                                                    0.9f, 0.999f, 1e-8f, c1, c2, 1, nullptr, loss_dev);
      ++steps;
    }
    MLP_TRY(hipGetLastError());
    float loss = 0.0f;
    MLP_TRY(hipMemcpyAsync(&loss, loss_dev, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MLP_TRY(hipStreamSynchronize(m->stream));
    if (losses) losses[e] = loss / (float)(steps ? steps : 1);
  }
  return MH_OK;
}

// Loss and gradient (w.r.t. the kParams weights) of the dataset rows
// [0, rows) in order, without updating: for gradient checks.
int mh_mlp_gradient(mh_mlp* m, float* grad, float* loss) {
  if (!m || m->rows == 0 || !grad) return MH_ERR_INVALID_ARGUMENT;
  const uint32_t rows = (uint32_t)m->rows;
  const uint32_t blocks = (rows + kChunk - 1) / kChunk;
  if (ensure_partials(m, blocks) != MH_OK) return MH_ERR_DEVICE;
  std::vector<uint32_t> order(rows);
  for (uint32_t i = 0; i < rows; ++i) order[i] = i;
  MLP_TRY(hipMemcpy(m->order, order.data(), rows * sizeof(uint32_t), hipMemcpyHostToDevice));
  MLP_TRY(hipMemsetAsync(m->scratch + kParams, 0, sizeof(float), m->stream));
  k_mlp_grad<<<blocks, 256, 0, m->stream>>>(m->net, m->feat, m->targ, m->order, 0, rows, 1.0f / (float)rows,
                                           m->partial);
  k_mlp_adam<<<(kParams + 1 + 255) / 256, 256, 0, m->stream>>>(m->net, m->m1, m->m2, m->partial, blocks, 0.0f, 0.9f,
                                                                0.999f, 1e-8f, 1.0f, 1.0f, 0, m->scratch,
                                                                m->scratch + kParams);
  MLP_TRY(hipStreamSynchronize(m->stream));
  MLP_TRY(hipMemcpy(grad, m->scratch, kParams * sizeof(float), hipMemcpyDeviceToHost));
  if (loss) MLP_TRY(hipMemcpy(loss, m->scratch + kParams, sizeof(float), hipMemcpyDeviceToHost));
  return MH_OK;
}

namespace {

// Device address of `scratch` on each device, looked up once per device with
// that device current.
constexpr int kMaxDevices = 64;
std::mutex scratch_lock;
float* scratch_ptr[kMaxDevices] = {};

int device_scratch(float** out) {
  int device = 0;
  MLP_TRY(hipGetDevice(&device));
  if (device < 0 || device >= kMaxDevices) return MH_ERR_UNSUPPORTED;
  std::lock_guard<std::mutex> guard(scratch_lock);
  if (!scratch_ptr[device]) {
    void* p = nullptr;
    MLP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(mh::scratch)));
    scratch_ptr[device] = static_cast<float*>(p);
  }
  *out = scratch_ptr[device];
  return MH_OK;
}

}  // namespace

// mh_mlp_gradient over caller-owned device rows, with no allocation: the block
// partials go through the file's static device window, kFoldBlocks blocks at a
// time, folded in block order into the caller's gradient and loss.
int mh_mlp_gradient_noalloc(mh_mlp* m, const float* features, const float* targets, size_t rows, float* gradient,
                            float* loss) {
  if (!m || !features || !targets || !gradient || !loss) return MH_ERR_INVALID_ARGUMENT;
  if (rows == 0 || rows > kMaxNoallocRows) return MH_ERR_INVALID_ARGUMENT;
  float* partial = nullptr;
  const int rc = device_scratch(&partial);
  if (rc != MH_OK) return rc;
  const uint32_t n = (uint32_t)rows;
  const uint32_t blocks = (n + kChunk - 1) / kChunk;
  const float inv_batch = 1.0f / (float)n;
  for (uint32_t b0 = 0; b0 < blocks; b0 += kFoldBlocks) {
    const uint32_t wb = blocks - b0 < kFoldBlocks ? blocks - b0 : kFoldBlocks;
    const uint32_t first = b0 * kChunk;
    // Rows of this window: the block at b0 starts at row `first`; the kernel
    // keeps the whole batch's row bound so the last block masks the same rows.
    k_mlp_grad<<<wb, 256, 0, m->stream>>>(m->net, features, targets, nullptr, first, n - first, inv_batch, partial);
    k_mlp_fold<<<(kParams + 1 + 255) / 256, 256, 0, m->stream>>>(partial, wb, b0 == 0 ? 1 : 0, gradient, loss);
  }
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipStreamSynchronize(m->stream));
  return MH_OK;
}

// BF16 MFMA inference of `rows` feature rows; out receives 3 floats per row.
int mh_mlp_predict(mh_mlp* m, const float* features, size_t rows, float* out) {
  if (!m || !features || !out) return MH_ERR_INVALID_ARGUMENT;
  if (rows == 0) return MH_OK;
  float *f = nullptr, *o = nullptr;
  MLP_TRY(hipMalloc(&f, rows * kFeatures * sizeof(float)));
  if (hipMalloc(&o, rows * kOut * sizeof(float)) != hipSuccess) {
    (void)hipFree(f);
    return MH_ERR_DEVICE;
  }
  int rc = MH_OK;
  if (hipMemcpy(f, features, rows * kFeatures * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) rc = MH_ERR_DEVICE;
  if (rc == MH_OK) {
    const uint32_t groups = (uint32_t)((rows + 15) / 16);
    uint32_t blocks = (groups + 3) / 4;
    if (blocks > 2048) blocks = 2048;
    k_mlp_predict<<<blocks, 256, 0, m->stream>>>(m->net, f, (uint32_t)rows, o);
    if (hipStreamSynchronize(m->stream) != hipSuccess ||
        hipMemcpy(out, o, rows * kOut * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      rc = MH_ERR_DEVICE;
  }
  (void)hipFree(f);
  (void)hipFree(o);
  return rc;
}

}  // extern "C"
