// ubp_learn.hip -- host side of the surrogate network: device model buffers,
// the BF16-MFMA forward / loss-gradient entry points and the trainer.
// Included by ubp_engine.hip (one translation unit).
#pragma once
#include "ubp_mlp.hip"

namespace ubp {

static_assert(offsetof(ubp_mlp_model, b1) == sizeof(float) * OFF_B1, "model layout");
static_assert(offsetof(ubp_mlp_model, w2) == sizeof(float) * OFF_W2, "model layout");
static_assert(offsetof(ubp_mlp_model, b3) == sizeof(float) * OFF_B3, "model layout");
static_assert(offsetof(ubp_mlp_model, w4) == sizeof(float) * OFF_W4, "model layout");
static_assert(offsetof(ubp_mlp_model, b4) == sizeof(float) * OFF_B4, "model layout");
static_assert(offsetof(ubp_mlp_model, in_mean) == sizeof(float) * MLP_NPARAM, "model layout");
static_assert(UBP_MLP_NPARAM == MLP_NPARAM, "parameter count");
constexpr int MLP_NSTAT = 2 * MLP_IN + 2 * MLP_OUT;

// Device copy of one model: FP32 parameters, packed BF16 fragments, stats.
struct MlpBuf {
  MlpBuf() = default;
  MlpBuf(const MlpBuf&) = delete;
  MlpBuf& operator=(const MlpBuf&) = delete;
  ~MlpBuf() { release(); }
  float* prm = nullptr;
  uint16_t* frag = nullptr;
  float* stat = nullptr;  // in_mean[32] in_inv_std[32] out_mean[3] out_std[3]

  void alloc() {
    if (prm) return;
    dalloc(&prm, MLP_NPARAM + 1);
    dalloc(&frag, FR_TOTAL);
    dalloc(&stat, MLP_NSTAT);
    HIPCHK(hipMemset(frag, 0, FR_TOTAL * 2));  // padding fragments stay zero
  }
  void release() { dfree(prm); dfree(frag); dfree(stat); }
  MlpDev view() const {
    return MlpDev{frag, prm, stat, stat + MLP_IN, stat + 2 * MLP_IN, stat + 2 * MLP_IN + MLP_OUT};
  }
  void upload(const ubp_mlp_model* m, hipStream_t st) {
    alloc();
    const float* src = reinterpret_cast<const float*>(m);
    HIPCHK(hipMemcpyAsync(prm, src, MLP_NPARAM * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(stat, src + MLP_NPARAM, MLP_NSTAT * 4, hipMemcpyHostToDevice, st));
    pack(st);
    HIPCHK(hipStreamSynchronize(st));
  }
  void pack(hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, (const float*)prm, frag);
    LAUNCHCHK("k_mlp_pack");
  }
  void download(ubp_mlp_model* m, hipStream_t st) const {
    float* dst = reinterpret_cast<float*>(m);
    HIPCHK(hipMemcpyAsync(dst, prm, MLP_NPARAM * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(dst + MLP_NPARAM, stat, MLP_NSTAT * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
};

static uint32_t mlp_fwd_grid(size_t rows) {
  return uint32_t(std::max<size_t>(1, std::min<size_t>((rows + 63) / 64, 2048)));
}

// Scoped HIP resources for the synchronous test entry points.
struct ScopedStream {
  hipStream_t s = nullptr;
  ScopedStream() { HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking)); }
  ~ScopedStream() { if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } }
};
template <class T> struct DevArr {
  T* p = nullptr;
  explicit DevArr(size_t n) { dalloc(&p, n); }
  ~DevArr() { dfree(p); }
};

static void require_gfx950() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    throw Error(UBP_ENODEV, std::string("device is ") + prop.gcnArchName + ", this build targets gfx950");
}

// One training step's gradient over `batch` rows idx[] of (X, T).
static void mlp_grad(const MlpBuf& M, const float* X, const float* T, const uint32_t* idx, uint32_t batch,
                     float* part, float* grad, hipStream_t st) {
  const uint32_t nblk = (batch + TR - 1) / TR;
  hipLaunchKernelGGL(k_mlp_train, dim3(nblk), dim3(256), 0, st, M.view(), X, T, idx, batch, part);
  LAUNCHCHK("k_mlp_train");
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, st, (const float*)part, nblk,
                     grad);
  LAUNCHCHK("k_mlp_reduce");
}

// Training data set + optimiser state on the device.
struct Trainer {
  hipStream_t st = nullptr;
  int dev = 0;
  size_t cap = 0, rows = 0;
  float* X = nullptr;   // [cap][32] raw features
  float* T = nullptr;   // [cap][4] raw targets (3 used), normalised at fit time
  float* Tn = nullptr;  // normalised targets
  MlpBuf M;
  float *m1 = nullptr, *m2 = nullptr, *grad = nullptr, *part = nullptr;
  uint32_t* idx = nullptr;
  uint32_t part_blk = 0, idx_cap = 0;
  uint64_t seed = 1;
  uint32_t step = 0;
  bool stats_set = false;

  ~Trainer() { release(); }
  void release() {
    if (st) (void)hipStreamSynchronize(st);
    dfree(X); dfree(T); dfree(Tn); dfree(m1); dfree(m2); dfree(grad); dfree(part); dfree(idx);
    M.release();
    if (st) { (void)hipStreamDestroy(st); st = nullptr; }
  }
  void create(size_t capacity, uint64_t sd) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) throw Error(UBP_ENODEV, "no HIP device visible");
    HIPCHK(hipGetDevice(&dev));
    require_gfx950();
    if (capacity == 0 || capacity > (size_t(1) << 31)) throw Error(UBP_EINVAL, "bad trainer capacity");
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    cap = capacity;
    seed = sd;
    dalloc(&X, cap * MLP_IN);
    dalloc(&T, cap * 4);
    dalloc(&Tn, cap * 4);
    M.alloc();
    dalloc(&m1, MLP_NPARAM);
    dalloc(&m2, MLP_NPARAM);
    dalloc(&grad, MLP_NPARAM + 1);
    HIPCHK(hipMemset(m1, 0, MLP_NPARAM * 4));
    HIPCHK(hipMemset(m2, 0, MLP_NPARAM * 4));
  }
};

}  // namespace ubp
