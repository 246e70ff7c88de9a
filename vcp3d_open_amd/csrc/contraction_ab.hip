// The forward sum-factorised interpolation of the entropy-variable mass-matrix
// solve (modal K^3 -> Gauss M^3 per variable, three 1D contractions;
// CG_MassMatrix.cpp:395-493 applies it, through the TensorProduct kernels,
// in every iteration) on synthetic data, in two forms, to measure which one
// gfx950 runs faster:
//   VALU: the volume kernel's form (one FP64 pencil per thread per stage, the
//         1D operator from the kernel arguments, LDS between stages);
//   MFMA: each stage as 16 x 16 x 4 FP64 matrix tiles
//         (__builtin_amdgcn_mfma_f64_16x16x4f64): rows the M outputs padded to
//         16, the K inputs in two k-steps of 4 (padded to 8), 16 pencils per
//         tile, tiles dealt round-robin to the four waves.
// One workgroup of 256 threads per element; each launch repeats the
// interpolation `inner` times from LDS, so the element's state is read from
// HBM once. K = 5, M = 7 (P = 4, rule 12).
#include <hip/hip_runtime.h>

#include <vector>

namespace {

constexpr int K = 5, M = 7, NT = 256;
constexpr int ND = K * K * K, NI = M * M * M;

struct AbOps {
  double A[M * K];   // row-major (output, input)
};

__device__ inline void sync() { __syncthreads(); }

// VALU form: z (items (v, j, i)), y (items (v, k', i), i slowest), x (rows)
__global__ __launch_bounds__(NT) void interp_valu(AbOps op, const double *in, double *out, int inner)
{
  __shared__ double sU[5 * ND], sZ[5 * M * K * K], sY[5 * M * M * K], sX[5 * NI];
  const int e = blockIdx.x, tid = threadIdx.x;
  for (int t = tid; t < 5 * ND; t += NT) sU[t] = in[(size_t) e * 5 * ND + t];
  sync();
  for (int rep = 0; rep < inner; ++rep) {
    for (int pc = tid; pc < 5 * K * K; pc += NT) {
      const int v = pc / (K * K), ji = pc - v * K * K;
      double x[K];
#pragma unroll
      for (int q = 0; q < K; ++q) x[q] = sU[(v * K + q) * K * K + ji];
#pragma unroll
      for (int n = 0; n < M; ++n) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < K; ++q) s += op.A[n * K + q] * x[q];
        sZ[(v * M + n) * K * K + ji] = s;
      }
    }
    sync();
    for (int pc = tid; pc < 5 * M * K; pc += NT) {
      const int i = pc / (5 * M), vk = pc - i * (5 * M);
      double x[K];
#pragma unroll
      for (int q = 0; q < K; ++q) x[q] = sZ[(vk * K + q) * K + i];
#pragma unroll
      for (int n = 0; n < M; ++n) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < K; ++q) s += op.A[n * K + q] * x[q];
        sY[(vk * M + n) * K + i] = s;
      }
    }
    sync();
    for (int pc = tid; pc < 5 * M * M; pc += NT) {
      double x[K];
#pragma unroll
      for (int q = 0; q < K; ++q) x[q] = sY[pc * K + q];
#pragma unroll
      for (int n = 0; n < M; ++n) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < K; ++q) s += op.A[n * K + q] * x[q];
        sX[pc * M + n] = s;
      }
    }
    sync();
  }
  for (int t = tid; t < 5 * NI; t += NT) out[(size_t) e * 5 * NI + t] = sX[t];
}

// One stage as MFMA tiles: Y[r][n] = sum_k Op[r][k] X[k][n], r < M, k < K,
// n < N; X and Y addressed through xa(k, n) and ya(r, n). Lane l holds
// A[l % 16][kstep * 4 + l / 16], B[kstep * 4 + l / 16][n0 + l % 16] and, in
// result register r, row l / 16 + 4 r of column n0 + l % 16 (the layout
// measured on gfx950).
template <int N, typename XA, typename YA>
__device__ inline void mfma_stage(const double a0, const double a1, const double *X, double *Y, XA xa, YA ya)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, kr = lane >> 4;
  constexpr int NTILE = (N + 15) / 16;
  for (int t = wave; t < NTILE; t += NT / 64) {
    const int n = t * 16 + col;
    const bool nin = n < N;
    const double b0 = nin ? X[xa(kr, n)] : 0.0;                      // k = kr
    const double b1 = (nin && 4 + kr < K) ? X[xa(4 + kr, n)] : 0.0;  // k = 4 + kr
    typedef double v4d __attribute__((ext_vector_type(4)));
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc, 0, 0, 0);
    if (nin)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = kr + 4 * r;
        if (row < M) Y[ya(row, n)] = acc[r];
      }
  }
}

__global__ __launch_bounds__(NT) void interp_mfma(AbOps op, const double *in, double *out, int inner)
{
  __shared__ double sU[5 * ND], sZ[5 * M * K * K], sY[5 * M * M * K], sX[5 * NI];
  const int e = blockIdx.x, tid = threadIdx.x, lane = tid & 63;
  for (int t = tid; t < 5 * ND; t += NT) sU[t] = in[(size_t) e * 5 * ND + t];
  // A operand of the two k-steps: row lane % 16, column kstep * 4 + lane / 16
  const int ar = lane & 15, ak = lane >> 4;
  const double a0 = (ar < M && ak < K) ? op.A[ar * K + ak] : 0.0;
  const double a1 = (ar < M && 4 + ak < K) ? op.A[ar * K + 4 + ak] : 0.0;
  sync();
  for (int rep = 0; rep < inner; ++rep) {
    // z: columns (v, j, i), 125
    mfma_stage<5 * K * K>(a0, a1, sU, sZ,
        [](int k, int n) { return (n / (K * K)) * ND + k * K * K + n % (K * K); },
        [](int r, int n) { return (n / (K * K)) * M * K * K + r * K * K + n % (K * K); });
    sync();
    // y: columns (v, k', i), 175
    mfma_stage<5 * M * K>(a0, a1, sZ, sY,
        [](int k, int n) { return (n / K) * K * K + k * K + n % K; },
        [](int r, int n) { return ((n / K) * M + r) * K + n % K; });
    sync();
    // x: columns (v, k', j'), 245
    mfma_stage<5 * M * M>(a0, a1, sY, sX,
        [](int k, int n) { return n * K + k; },
        [](int r, int n) { return n * M + r; });
    sync();
  }
  for (int t = tid; t < 5 * NI; t += NT) out[(size_t) e * 5 * NI + t] = sX[t];
}

}  // namespace

// impl 0: VALU, 1: MFMA. Interpolates n_elem synthetic elements `inner`
// times per launch, `reps` timed launches; ms = mean launch time, out_sum =
// sum of the outputs of the last launch (for comparing the two forms).
extern "C" int vcp_ab_interpolation(int impl, int n_elem, int inner, int reps, double *ms, double *out_sum)
{
  if (n_elem <= 0 || inner <= 0 || reps <= 0 || !ms || !out_sum) return -1;
  AbOps op;
  for (int n = 0; n < M; ++n)
    for (int q = 0; q < K; ++q) op.A[n * K + q] = 0.5 + 0.25 * ((n * 3 + q * 7) % 11) / 11.0;
  std::vector<double> h((size_t) n_elem * 5 * ND);
  for (size_t i = 0; i < h.size(); ++i) h[i] = 1.0 + 0.001 * (double) (i % 997);
  double *dIn = nullptr, *dOut = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = -2;
  if (hipMalloc((void **) &dIn, h.size() * sizeof(double)) != hipSuccess) goto done;
  if (hipMalloc((void **) &dOut, (size_t) n_elem * 5 * NI * sizeof(double)) != hipSuccess) goto done;
  if (hipMemcpy(dIn, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) goto done;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) goto done;
  for (int r = -1; r < reps; ++r) {                    // one untimed warm-up launch
    if (r == 0 && hipEventRecord(e0, 0) != hipSuccess) goto done;
    if (impl == 0) hipLaunchKernelGGL(interp_valu, dim3(n_elem), dim3(NT), 0, 0, op, dIn, dOut, inner);
    else hipLaunchKernelGGL(interp_mfma, dim3(n_elem), dim3(NT), 0, 0, op, dIn, dOut, inner);
    if (hipGetLastError() != hipSuccess) goto done;
  }
  if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) goto done;
  {
    float t = 0.f;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) goto done;
    *ms = t / reps;
    std::vector<double> o((size_t) n_elem * 5 * NI);
    if (hipMemcpy(o.data(), dOut, o.size() * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) goto done;
    double s = 0.0;
    for (double v : o) s += v;
    *out_sum = s;
  }
  rc = 0;
done:
  if (e0) (void) hipEventDestroy(e0);
  if (e1) (void) hipEventDestroy(e1);
  if (dIn) (void) hipFree(dIn);
  if (dOut) (void) hipFree(dOut);
  return rc;
}
