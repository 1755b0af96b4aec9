// Sweep surrogate MLP for gfx950: BF16 MFMA GEMMs with FP32 accumulation, GELU, sigmoid,
// mean squared error and Adam, all as hand-written HIP kernels.  C ABI: include/mbg_mlp.h.
//
// GEMM: C[M][N] (FP32) = A[M][K] * B[K][N] over BF16 operands, 64x64 block tiles staged
// through LDS, four waves each computing 32x32 with 2x2 v_mfma_f32_16x16x32_bf16 tiles.
// Operand fragments of the 16x16x32 MFMA: lane l supplies A[l % 16][8 (l / 16) .. +7] and
// B[8 (l / 16) .. +7][l % 16]; it receives D[4 (l / 16) + q][l % 16], q = 0..3.  Transposed
// operands are transposed while staging into LDS, so one kernel serves the forward pass
// (X W^T), the input gradient (dY W) and the weight gradient (dY^T X).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/mbg_mlp.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int TK = 32, LDT = TK + 8;

// A(m, k) = AK ? A[m * lda + k] : A[k * lda + m];  B(k, n) = BK ? B[n * ldb + k] : B[k * ldb + n].
// Block tile BM x BN (64 or 128 each), four waves in 2 x 2, each wave (BM/2) x (BN/2) of
// 16 x 16 MFMA tiles.  The next K step's operands are loaded into registers while the
// current step's MFMAs run from LDS.
template <int ROWS>
struct Stage {
    bf16x8 v[ROWS / 64];
};

// A ROWS x 32 operand tile: with KCONTIG the k index is contiguous in memory (row-major
// [row][k]), otherwise the row index is ([k][row]).  Each thread moves ROWS/64 groups of 8.
template <bool KCONTIG, int ROWS>
__device__ __forceinline__ void stage_load(Stage<ROWS>& st, const bf16* __restrict__ X, long ld, long r0, int k0, int tid) {
#pragma unroll
    for (int p = 0; p < ROWS / 64; p++) {
        if (KCONTIG) {
            const int r = p * 64 + (tid >> 2), kk = (tid & 3) * 8;
            st.v[p] = *(const bf16x8*)(X + (r0 + r) * ld + k0 + kk);
        } else {
            const int idx = p * 256 + tid;
            const int kq = idx / (ROWS / 8), mq = (idx % (ROWS / 8)) * 8;
            st.v[p] = *(const bf16x8*)(X + (long)(k0 + kq) * ld + r0 + mq);
        }
    }
}

template <bool KCONTIG, int ROWS>
__device__ __forceinline__ void stage_store(const Stage<ROWS>& st, bf16* __restrict__ S, int tid) {
#pragma unroll
    for (int p = 0; p < ROWS / 64; p++) {
        if (KCONTIG) {
            const int r = p * 64 + (tid >> 2), kk = (tid & 3) * 8;
            *(bf16x8*)(S + r * LDT + kk) = st.v[p];
        } else {
            const int idx = p * 256 + tid;
            const int kq = idx / (ROWS / 8), mq = (idx % (ROWS / 8)) * 8;
#pragma unroll
            for (int i = 0; i < 8; i++) S[(mq + i) * LDT + kq] = st.v[p][i];
        }
    }
}

template <bool AK, bool BK, int BM, int BN>
__global__ __launch_bounds__(256) void gemm_bf16(const bf16* __restrict__ A, long lda, const bf16* __restrict__ B,
                                                 long ldb, float* __restrict__ C, long ldc, int K, int accumulate) {
    __shared__ bf16 As[BM * LDT];  // [m][k]
    __shared__ bf16 Bs[BN * LDT];  // [n][k]
    constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long m0 = (long)blockIdx.y * BM, n0 = (long)blockIdx.x * BN;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
    f32x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    Stage<BM> sa;
    Stage<BN> sb;
    stage_load<AK, BM>(sa, A, lda, m0, 0, tid);
    stage_load<BK, BN>(sb, B, ldb, n0, 0, tid);
    for (int k0 = 0; k0 < K; k0 += TK) {
        stage_store<AK, BM>(sa, As, tid);
        stage_store<BK, BN>(sb, Bs, tid);
        __syncthreads();
        if (k0 + TK < K) {
            stage_load<AK, BM>(sa, A, lda, m0, k0 + TK, tid);
            stage_load<BK, BN>(sb, B, ldb, n0, k0 + TK, tid);
        }
        bf16x8 a[FM], b[FN];
#pragma unroll
        for (int i = 0; i < FM; i++) a[i] = *(const bf16x8*)(As + (wm + i * 16 + (lane & 15)) * LDT + (lane >> 4) * 8);
#pragma unroll
        for (int j = 0; j < FN; j++) b[j] = *(const bf16x8*)(Bs + (wn + j * 16 + (lane & 15)) * LDT + (lane >> 4) * 8);
#pragma unroll
        for (int i = 0; i < FM; i++)
#pragma unroll
            for (int j = 0; j < FN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++)
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const long row = m0 + wm + i * 16 + (lane >> 4) * 4 + q;
                const long col = n0 + wn + j * 16 + (lane & 15);
                float* p = C + row * ldc + col;
                *p = (accumulate ? *p : 0.f) + acc[i][j][q];
            }
}

__device__ __forceinline__ float gelu(float z, float* deriv) {
    const float c = 0.7978845608028654f, a = 0.044715f;
    const float u = c * (z + a * z * z * z);
    const float t = tanhf(u);
    if (deriv) *deriv = 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * c * (1.f + 3.f * a * z * z);
    return 0.5f * z * (1.f + t);
}

// Z = C + b (kept for the backward pass); H = GELU(Z) as BF16.
__global__ void bias_gelu(const float* __restrict__ Cm, const float* __restrict__ bias, float* __restrict__ Z,
                          bf16* __restrict__ H, long rows, int cols) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= rows * cols) return;
    const float z = Cm[i] + bias[i % cols];
    Z[i] = z;
    H[i] = (bf16)gelu(z, nullptr);
}

// Output layer: yhat = sigmoid(C + b); with targets, dZ = 2 (yhat - y) yhat (1 - yhat) / (rows cols)
// as BF16 and FP32, and the squared errors.
__global__ void sigmoid_mse(const float* __restrict__ Cm, const float* __restrict__ bias, const float* __restrict__ Y,
                            float* __restrict__ Yhat, float* __restrict__ dZ, bf16* __restrict__ dZb,
                            float* __restrict__ sq, long rows, int cols, float scale) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= rows * cols) return;
    const float z = Cm[i] + bias[i % cols];
    const float yh = 1.f / (1.f + expf(-z));
    Yhat[i] = yh;
    if (Y) {
        const float e = yh - Y[i];
        const float d = 2.f * e * yh * (1.f - yh) * scale;
        dZ[i] = d;
        dZb[i] = (bf16)d;
        sq[i] = e * e;
    }
}

// dZ = dH * GELU'(Z) (dH from the GEMM, FP32), as BF16 and FP32.
__global__ void gelu_backward(const float* __restrict__ dH, const float* __restrict__ Z, float* __restrict__ dZ,
                              bf16* __restrict__ dZb, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float d;
    gelu(Z[i], &d);
    const float v = dH[i] * d;
    dZ[i] = v;
    dZb[i] = (bf16)v;
}

// out[c] = sum over rows of X[r][c] (bias gradients, loss sums).  A block of 1,024 threads
// takes 64 columns: its 16 waves each sum every 16th row of them (a wave reads 64 consecutive
// floats per row), then wave 0 adds the 16 partial sums from LDS in a fixed order, so the
// result does not depend on scheduling.
constexpr int kColTile = 64, kRowGroups = 16;
__global__ __launch_bounds__(1024) void column_sums(const float* __restrict__ X, float* __restrict__ out, long rows,
                                                    int cols) {
    __shared__ float part[kRowGroups][kColTile];
    const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
    const int c = blockIdx.x * kColTile + lane;
    float s = 0.f;
    if (c < cols) {
#pragma unroll 4
        for (long r = grp; r < rows; r += kRowGroups) s += X[r * cols + c];
    }
    part[grp][lane] = s;
    __syncthreads();
    if (grp == 0 && c < cols) {
        float t = 0.f;
        for (int k = 0; k < kRowGroups; k++) t += part[k][lane];
        out[c] = t;
    }
}
unsigned col_blocks(int cols) { return (unsigned)((cols + kColTile - 1) / kColTile); }

// Gather the rows idx[r] of the device training set into a batch (BF16 inputs, FP32 targets).
__global__ void gather_rows(const bf16* __restrict__ Xs, const float* __restrict__ Ys, const uint32_t* __restrict__ idx,
                            bf16* __restrict__ Xb, float* __restrict__ Yb, long rows, int in_p, int out) {
    const long r = blockIdx.y;
    const long src = idx[r];
    for (int c = blockIdx.x * 256 + threadIdx.x; c < in_p; c += gridDim.x * 256) Xb[r * in_p + c] = Xs[src * in_p + c];
    for (int c = blockIdx.x * 256 + threadIdx.x; c < out; c += gridDim.x * 256) Yb[r * out + c] = Ys[src * out + c];
}

// Adam on FP32 master parameters, refreshing their BF16 copy.
__global__ void adam(float* __restrict__ w, bf16* __restrict__ wb, const float* __restrict__ g, float* __restrict__ m,
                     float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = (bf16)wi;
}

__global__ void to_bf16(const float* __restrict__ x, bf16* __restrict__ y, long n) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = (bf16)x[i];
}

// H[r][c] = BF16(X[r][c]) for r < rows, c < n_in; zero in the padding columns and in rows
// rows..padded_rows-1.
__global__ void pad_bf16(const float* __restrict__ X, bf16* __restrict__ H, long rows, long padded_rows, int n_in,
                         int in_p) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= padded_rows * in_p) return;
    const long r = i / in_p;
    const int c = (int)(i - r * in_p);
    H[i] = (r < rows && c < n_in) ? (bf16)X[r * n_in + c] : (bf16)0.f;
}

// One block per row: mean over `count` members' predictions P[k][row][c], and the row's largest
// population standard deviation over members of any output.
__global__ __launch_bounds__(256) void ensemble_reduce(const float* __restrict__ P, int count, long n, int n_out,
                                                       float* __restrict__ mean, float* __restrict__ spread) {
    const long row = blockIdx.x;
    __shared__ float red[256];
    float worst = 0.f;
    for (int c = threadIdx.x; c < n_out; c += 256) {
        float s = 0.f;
        for (int k = 0; k < count; k++) s += P[((long)k * n + row) * n_out + c];
        const float mu = s / (float)count;
        float v = 0.f;
        for (int k = 0; k < count; k++) {
            const float d = P[((long)k * n + row) * n_out + c] - mu;
            v += d * d;
        }
        mean[row * n_out + c] = mu;
        worst = fmaxf(worst, sqrtf(v / (float)count));
    }
    red[threadIdx.x] = worst;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && spread) spread[row] = red[0];
}

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(4, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    ~Buf() {
        if (p) (void)hipFree(p);
    }
    void alloc(size_t count) {
        if (p && count <= n) return;
        if (p) (void)hipFree(p);
        p = nullptr;
        check(hipMalloc((void**)&p, std::max<size_t>(1, count) * sizeof(T)), "hipMalloc");
        n = count;
    }
};

unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

// Scratch of mbg_mlp_ensemble_predict_noalloc: the members' outputs of one chunk of rows,
// [member][row][output].  64 MiB per device, held by every device the library runs on until the
// process ends.  Host code uses only the address hipGetSymbolAddress gives for it on each
// device (scratch_addr), never the array's name.
constexpr size_t kScratchFloats = 16u << 20;
static __device__ __align__(256) float scratch[kScratchFloats];

std::mutex g_scratch_mu;
std::map<int, float*> g_scratch_addr;  // device -> that device's copy of scratch

// Looks the array up once per device, with that device current (from mbg_mlp_create).
void init_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_addr.count(device)) return;
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw Error(4, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch_addr[device] = static_cast<float*>(p);
}

float* scratch_addr(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    const auto it = g_scratch_addr.find(device);
    return it == g_scratch_addr.end() ? nullptr : it->second;
}

}  // namespace

struct mbg_mlp {
    int device = 0;
    hipStream_t s = nullptr;
    std::string err;
    int n_in = 0, in_p = 0, hidden = 0, n_hidden = 0, n_out = 0, max_batch = 0;
    std::vector<int> din, dout;            // per layer
    std::vector<size_t> woff, boff;        // offsets into the parameter vector
    size_t nparams = 0;
    Buf<float> w, g, m1, m2;               // FP32 master parameters, gradients, Adam moments
    Buf<bf16> wb;                          // BF16 copy of the weights
    std::vector<Buf<float>> Z, Cm;         // per layer pre-activations, GEMM outputs
    std::vector<Buf<bf16>> H;              // per layer BF16 activations (H[0] = input batch)
    Buf<float> Y, Yhat, dZ, dH, sq, sums;
    Buf<bf16> dZb;
    Buf<bf16> Xset;                        // training set
    Buf<float> Yset;
    Buf<uint32_t> idx;
    Buf<float> Xin, Pens, Mean, Spread;    // ensemble prediction: inputs, member outputs, results
    size_t nset = 0;
    long step = 0;

    template <int BM, int BN>
    void launch_gemm(const bf16* A, long lda, bool ak, const bf16* B, long ldb, bool bk, float* C, long ldc, long M,
                     long N, long K, bool acc, hipStream_t st) {
        dim3 grid((unsigned)(N / BN), (unsigned)(M / BM));
        if (ak && bk) gemm_bf16<true, true, BM, BN><<<grid, 256, 0, st>>>(A, lda, B, ldb, C, ldc, (int)K, acc);
        else if (ak && !bk) gemm_bf16<true, false, BM, BN><<<grid, 256, 0, st>>>(A, lda, B, ldb, C, ldc, (int)K, acc);
        else if (!ak && bk) gemm_bf16<false, true, BM, BN><<<grid, 256, 0, st>>>(A, lda, B, ldb, C, ldc, (int)K, acc);
        else gemm_bf16<false, false, BM, BN><<<grid, 256, 0, st>>>(A, lda, B, ldb, C, ldc, (int)K, acc);
    }

    // C = A B on stream st (the network's own stream when null)
    void gemm(const bf16* A, long lda, bool ak, const bf16* B, long ldb, bool bk, float* C, long ldc, long M, long N,
              long K, bool acc, hipStream_t st = nullptr) {
        if (M % 64 || N % 64 || K % TK) throw Error(3, "GEMM dimensions must be multiples of 64x64x32");
        if (!st) st = s;
        // the largest tile that still gives every CU a block
        const bool bm = M % 128 == 0 && (M / 128) * (N / 64) >= 256;
        const bool bn = N % 128 == 0 && (M / (bm ? 128 : 64)) * (N / 128) >= 256;
        if (bm && bn) launch_gemm<128, 128>(A, lda, ak, B, ldb, bk, C, ldc, M, N, K, acc, st);
        else if (bm) launch_gemm<128, 64>(A, lda, ak, B, ldb, bk, C, ldc, M, N, K, acc, st);
        else launch_gemm<64, 64>(A, lda, ak, B, ldb, bk, C, ldc, M, N, K, acc, st);
        check(hipGetLastError(), "gemm launch");
    }

    // Forward pass of `rows` rows already in H[0]; targets in Y when with_loss.  On stream st
    // (the network's own stream when null).
    void forward(long rows, bool with_loss, hipStream_t st = nullptr) {
        if (!st) st = s;
        const int L = (int)din.size();
        for (int l = 0; l < L; l++) {
            gemm(H[l].p, din[l], true, wb.p + woff[l], din[l], true, Cm[l].p, dout[l], rows, dout[l], din[l], false, st);
            if (l < L - 1) {
                bias_gelu<<<blocks(rows * dout[l]), 256, 0, st>>>(Cm[l].p, w.p + boff[l], Z[l].p, H[l + 1].p, rows,
                                                                 dout[l]);
            } else {
                sigmoid_mse<<<blocks(rows * dout[l]), 256, 0, st>>>(Cm[l].p, w.p + boff[l], with_loss ? Y.p : nullptr,
                                                                    Yhat.p, dZ.p, dZb.p, sq.p, rows, dout[l],
                                                                    1.f / (float)(rows * dout[l]));
            }
        }
    }

    // Backward pass after forward(rows, true): gradients into g.
    void backward(long rows) {
        const int L = (int)din.size();
        for (int l = L - 1; l >= 0; l--) {
            // weight gradient dW[out][in] = dZ^T H_l ; bias gradient = column sums of dZ
            gemm(dZb.p, dout[l], false, H[l].p, din[l], false, g.p + woff[l], din[l], dout[l], din[l], rows, false);
            column_sums<<<col_blocks(dout[l]), 1024, 0, s>>>(dZ.p, g.p + boff[l], rows, dout[l]);
            if (l > 0) {
                // dH = dZ W (rows x din) ; dZ_{l-1} = dH * GELU'(Z_{l-1})
                gemm(dZb.p, dout[l], true, wb.p + woff[l], din[l], false, dH.p, din[l], rows, din[l], dout[l], false);
                gelu_backward<<<blocks(rows * din[l]), 256, 0, s>>>(dH.p, Z[l - 1].p, dZ.p, dZb.p, rows * din[l]);
            }
        }
    }

    float loss(long rows) {
        column_sums<<<col_blocks(n_out), 1024, 0, s>>>(sq.p, sums.p, rows, n_out);
        std::vector<float> h(n_out);
        check(hipMemcpyAsync(h.data(), sums.p, n_out * 4, hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
        check(hipStreamSynchronize(s), "loss");
        double t = 0;
        for (float v : h) t += v;
        return (float)(t / ((double)rows * n_out));
    }

    void upload_batch(const float* x, const float* y, long rows) {
        std::vector<bf16> xb((size_t)rows * in_p, (bf16)0.f);
        for (long r = 0; r < rows; r++)
            for (int c = 0; c < n_in; c++) xb[r * in_p + c] = (bf16)x[r * n_in + c];
        check(hipMemcpyAsync(H[0].p, xb.data(), xb.size() * 2, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        if (y) check(hipMemcpyAsync(Y.p, y, (size_t)rows * n_out * 4, hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        check(hipStreamSynchronize(s), "upload");
    }

    void adam_step(float lr) {
        step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
        for (size_t l = 0; l < din.size(); l++) {
            const long nw = (long)din[l] * dout[l];
            adam<<<blocks(nw), 256, 0, s>>>(w.p + woff[l], wb.p + woff[l], g.p + woff[l], m1.p + woff[l], m2.p + woff[l], nw,
                                            lr, b1, b2, 1e-8f, c1, c2);
            adam<<<blocks(dout[l]), 256, 0, s>>>(w.p + boff[l], nullptr, g.p + boff[l], m1.p + boff[l], m2.p + boff[l],
                                                 dout[l], lr, b1, b2, 1e-8f, c1, c2);
        }
    }

    void refresh_bf16() {
        for (size_t l = 0; l < din.size(); l++)
            to_bf16<<<blocks((long)din[l] * dout[l]), 256, 0, s>>>(w.p + woff[l], wb.p + woff[l], (long)din[l] * dout[l]);
    }
};

namespace {

template <class F>
mbg_status guarded(mbg_mlp* m, F&& f) {
    try {
        check(hipSetDevice(m->device), "hipSetDevice");
        f();
        m->err.clear();
        return 0;
    } catch (const Error& e) {
        m->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        m->err = e.what();
        return 3;
    }
}

}  // namespace

extern "C" {

mbg_status mbg_mlp_create(int device, int n_in, int hidden, int hidden_layers, int n_out, int max_batch, uint64_t seed,
                          mbg_mlp** out) {
    if (!out || n_in <= 0 || hidden <= 0 || hidden % 64 || hidden_layers < 1 || n_out <= 0 || n_out % 64 ||
        max_batch <= 0 || max_batch % 64)
        return 3;
    *out = nullptr;
    auto m = new mbg_mlp();
    m->device = device;
    m->n_in = n_in;
    m->in_p = (n_in + 63) / 64 * 64;
    m->hidden = hidden;
    m->n_hidden = hidden_layers;
    m->n_out = n_out;
    m->max_batch = max_batch;
    const mbg_status st = guarded(m, [&] {
        init_scratch(device);
        check(hipStreamCreate(&m->s), "hipStreamCreate");
        m->din.push_back(m->in_p);
        for (int l = 0; l < hidden_layers; l++) {
            m->dout.push_back(hidden);
            m->din.push_back(hidden);
        }
        m->dout.push_back(n_out);
        for (size_t l = 0; l < m->din.size(); l++) {
            m->woff.push_back(m->nparams);
            m->nparams += (size_t)m->din[l] * m->dout[l];
            m->boff.push_back(m->nparams);
            m->nparams += m->dout[l];
        }
        std::vector<float> init(m->nparams, 0.f);
        std::mt19937_64 rng(seed);
        for (size_t l = 0; l < m->din.size(); l++) {
            const int fan_in = l == 0 ? n_in : m->din[l];
            const float lim = std::sqrt(6.f / (float)(fan_in + m->dout[l]));
            std::uniform_real_distribution<float> u(-lim, lim);
            for (int o = 0; o < m->dout[l]; o++)
                for (int i = 0; i < m->din[l]; i++)
                    init[m->woff[l] + (size_t)o * m->din[l] + i] = (l == 0 && i >= n_in) ? 0.f : u(rng);
        }
        m->w.alloc(m->nparams);
        m->g.alloc(m->nparams);
        m->m1.alloc(m->nparams);
        m->m2.alloc(m->nparams);
        m->wb.alloc(m->nparams);
        check(hipMemcpy(m->w.p, init.data(), m->nparams * 4, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemset(m->m1.p, 0, m->nparams * 4), "hipMemset");
        check(hipMemset(m->m2.p, 0, m->nparams * 4), "hipMemset");
        check(hipMemset(m->wb.p, 0, m->nparams * 2), "hipMemset");
        m->refresh_bf16();
        const int L = (int)m->din.size();
        m->Z.resize(L);
        m->Cm.resize(L);
        m->H.resize(L);
        for (int l = 0; l < L; l++) {
            m->Z[l].alloc((size_t)max_batch * m->dout[l]);
            m->Cm[l].alloc((size_t)max_batch * m->dout[l]);
            m->H[l].alloc((size_t)max_batch * m->din[l]);
        }
        const int widest = std::max({m->in_p, hidden, n_out});
        m->Y.alloc((size_t)max_batch * n_out);
        m->Yhat.alloc((size_t)max_batch * n_out);
        m->sq.alloc((size_t)max_batch * n_out);
        m->dZ.alloc((size_t)max_batch * widest);
        m->dZb.alloc((size_t)max_batch * widest);
        m->dH.alloc((size_t)max_batch * widest);
        m->sums.alloc(widest);
        check(hipStreamSynchronize(m->s), "create");
    });
    if (st) {
        delete m;
        return st;
    }
    *out = m;
    return 0;
}

void mbg_mlp_destroy(mbg_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->s) (void)hipStreamDestroy(m->s);
    delete m;
}

const char* mbg_mlp_last_error(const mbg_mlp* m) { return m ? m->err.c_str() : "null"; }
size_t mbg_mlp_num_params(const mbg_mlp* m) { return m ? m->nparams : 0; }
int mbg_mlp_padded_inputs(const mbg_mlp* m) { return m ? m->in_p : 0; }

mbg_status mbg_mlp_get_params(mbg_mlp* m, float* out) {
    if (!m || !out) return 3;
    return guarded(m, [&] { check(hipMemcpy(out, m->w.p, m->nparams * 4, hipMemcpyDeviceToHost), "hipMemcpy"); });
}

mbg_status mbg_mlp_set_params(mbg_mlp* m, const float* in) {
    if (!m || !in) return 3;
    return guarded(m, [&] {
        check(hipMemcpy(m->w.p, in, m->nparams * 4, hipMemcpyHostToDevice), "hipMemcpy");
        m->refresh_bf16();
        check(hipStreamSynchronize(m->s), "set_params");
    });
}

mbg_status mbg_mlp_gradients(mbg_mlp* m, const float* x, const float* y, int batch, float* loss, float* grads) {
    if (!m || !x || !y || batch <= 0 || batch % 64 || batch > m->max_batch) return 3;
    return guarded(m, [&] {
        m->upload_batch(x, y, batch);
        m->forward(batch, true);
        m->backward(batch);
        if (loss) *loss = m->loss(batch);
        if (grads) check(hipMemcpy(grads, m->g.p, m->nparams * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    });
}

mbg_status mbg_mlp_load_dataset(mbg_mlp* m, const float* x, const float* y, size_t n) {
    if (!m || !x || !y || n == 0) return 3;
    return guarded(m, [&] {
        m->Xin.alloc(n * m->n_in);
        m->Xset.alloc(n * m->in_p);
        m->Yset.alloc(n * m->n_out);
        check(hipMemcpy(m->Xin.p, x, n * m->n_in * 4, hipMemcpyHostToDevice), "hipMemcpy");
        check(hipMemcpy(m->Yset.p, y, n * m->n_out * 4, hipMemcpyHostToDevice), "hipMemcpy");
        pad_bf16<<<blocks((long)n * m->in_p), 256, 0, m->s>>>(m->Xin.p, m->Xset.p, (long)n, (long)n, m->n_in, m->in_p);
        check(hipGetLastError(), "pad_bf16 launch");
        check(hipStreamSynchronize(m->s), "load_dataset");
        m->nset = n;
        m->idx.alloc(n);
    });
}

mbg_status mbg_mlp_train(mbg_mlp* m, int epochs, int batch, float lr, uint64_t shuffle_seed, float* last_loss) {
    if (!m || epochs <= 0 || batch <= 0 || batch % 64 || batch > m->max_batch) return 3;
    if (m->nset < (size_t)batch) return 3;
    return guarded(m, [&] {
        std::mt19937_64 rng(shuffle_seed);
        std::vector<uint32_t> perm(m->nset);
        for (size_t i = 0; i < perm.size(); i++) perm[i] = (uint32_t)i;
        const size_t steps = m->nset / batch;
        double sum = 0.0;
        for (int e = 0; e < epochs; e++) {
            // the previous epoch's upload of perm (pageable memory, possibly still pending)
            // must finish before perm is shuffled again
            if (e > 0) check(hipStreamSynchronize(m->s), "epoch");
            std::shuffle(perm.begin(), perm.end(), rng);
            check(hipMemcpyAsync(m->idx.p, perm.data(), perm.size() * 4, hipMemcpyHostToDevice, m->s), "hipMemcpyAsync");
            sum = 0.0;
            for (size_t st = 0; st < steps; st++) {
                dim3 grid(4, (unsigned)batch);
                gather_rows<<<grid, 256, 0, m->s>>>(m->Xset.p, m->Yset.p, m->idx.p + st * batch, m->H[0].p, m->Y.p, batch,
                                                    m->in_p, m->n_out);
                m->forward(batch, true);
                m->backward(batch);
                m->adam_step(lr);
                if (e == epochs - 1) sum += m->loss(batch);
            }
        }
        check(hipStreamSynchronize(m->s), "train");
        if (last_loss) *last_loss = (float)(sum / (double)steps);
    });
}

mbg_status mbg_mlp_ensemble_predict(mbg_mlp* const* members, int count, const float* x, size_t n, float* mean,
                                    float* spread, double* compute_ms) {
    if (!members || count <= 0 || !members[0] || !x || !mean || n == 0) return 3;
    mbg_mlp* m0 = members[0];
    for (int k = 0; k < count; k++) {
        const mbg_mlp* mk = members[k];
        if (!mk || mk->device != m0->device || mk->n_in != m0->n_in || mk->n_out != m0->n_out) {
            m0->err = "ensemble members must share the device and the input and output widths";
            return 3;
        }
    }
    return guarded(m0, [&] {
        const int n_in = m0->n_in, n_out = m0->n_out;
        m0->Xin.alloc(n * n_in);
        m0->Pens.alloc((size_t)count * n * n_out);
        m0->Mean.alloc(n * n_out);
        m0->Spread.alloc(n);
        check(hipMemcpyAsync(m0->Xin.p, x, n * n_in * 4, hipMemcpyHostToDevice, m0->s), "hipMemcpyAsync");
        std::vector<hipEvent_t> ev(count + 2);
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
        check(hipEventRecord(ev[0], m0->s), "hipEventRecord");
        for (int k = 0; k < count; k++) {
            mbg_mlp* mk = members[k];
            check(hipStreamWaitEvent(mk->s, ev[0], 0), "hipStreamWaitEvent");
            for (size_t r0 = 0; r0 < n; r0 += mk->max_batch) {
                const long rows = (long)std::min<size_t>(mk->max_batch, n - r0);
                const long padded = (rows + 63) / 64 * 64;
                pad_bf16<<<blocks(padded * mk->in_p), 256, 0, mk->s>>>(m0->Xin.p + r0 * n_in, mk->H[0].p, rows, padded,
                                                                      n_in, mk->in_p);
                mk->forward(padded, false);
                check(hipMemcpyAsync(m0->Pens.p + ((size_t)k * n + r0) * n_out, mk->Yhat.p, (size_t)rows * n_out * 4,
                                     hipMemcpyDeviceToDevice, mk->s),
                      "hipMemcpyAsync");
            }
            check(hipEventRecord(ev[k + 1], mk->s), "hipEventRecord");
            check(hipStreamWaitEvent(m0->s, ev[k + 1], 0), "hipStreamWaitEvent");
        }
        ensemble_reduce<<<(unsigned)n, 256, 0, m0->s>>>(m0->Pens.p, count, (long)n, n_out, m0->Mean.p, m0->Spread.p);
        check(hipGetLastError(), "ensemble_reduce launch");
        check(hipEventRecord(ev[count + 1], m0->s), "hipEventRecord");
        check(hipMemcpyAsync(mean, m0->Mean.p, n * n_out * 4, hipMemcpyDeviceToHost, m0->s), "hipMemcpyAsync");
        if (spread) check(hipMemcpyAsync(spread, m0->Spread.p, n * 4, hipMemcpyDeviceToHost, m0->s), "hipMemcpyAsync");
        check(hipStreamSynchronize(m0->s), "ensemble predict");
        float t = 0.f;
        check(hipEventElapsedTime(&t, ev[0], ev[count + 1]), "hipEventElapsedTime");
        for (auto& e : ev) (void)hipEventDestroy(e);
        if (compute_ms) *compute_ms = t;
    });
}

mbg_status mbg_mlp_ensemble_predict_noalloc(mbg_mlp* const* members, int count, const float* x, size_t n, float* mean,
                                            float* spread) {
    if (!members || count <= 0 || !members[0] || !x || !mean || n == 0) return 3;
    mbg_mlp* m0 = members[0];
    long chunk = (long)m0->max_batch;
    for (int k = 0; k < count; k++) {
        const mbg_mlp* mk = members[k];
        if (!mk || mk->device != m0->device || mk->n_in != m0->n_in || mk->n_out != m0->n_out) {
            m0->err = "ensemble members must share the device and the input and output widths";
            return 3;
        }
        chunk = std::min(chunk, (long)mk->max_batch);
    }
    // rows per chunk: every member's batch and the scratch's capacity, in whole blocks of 64
    const size_t per_row = (size_t)count * (size_t)m0->n_out;
    const long cap_rows = (long)(kScratchFloats / per_row) / 64 * 64;
    if (cap_rows < 64) {
        m0->err = "ensemble too large for the no-allocation scratch: members x outputs x 64 rows must fit " +
                  std::to_string(kScratchFloats) + " floats (" + std::to_string(count) + " x " +
                  std::to_string(m0->n_out) + " given)";
        return 3;
    }
    chunk = std::min(chunk, cap_rows);
    return guarded(m0, [&] {
        float* P = scratch_addr(m0->device);
        if (!P) throw Error(3, "no scratch address for this device (no network was created on it)");
        const int n_in = m0->n_in, n_out = m0->n_out;
        hipStream_t st = m0->s;
        for (size_t r0 = 0; r0 < n; r0 += (size_t)chunk) {
            const long rows = (long)std::min<size_t>((size_t)chunk, n - r0);
            const long padded = (rows + 63) / 64 * 64;
            for (int k = 0; k < count; k++) {
                mbg_mlp* mk = members[k];
                pad_bf16<<<blocks(padded * mk->in_p), 256, 0, st>>>(x + r0 * n_in, mk->H[0].p, rows, padded, n_in,
                                                                   mk->in_p);
                mk->forward(padded, false, st);
                check(hipMemcpyAsync(P + (size_t)k * rows * n_out, mk->Yhat.p, (size_t)rows * n_out * 4,
                                     hipMemcpyDeviceToDevice, st),
                      "hipMemcpyAsync");
            }
            ensemble_reduce<<<(unsigned)rows, 256, 0, st>>>(P, count, rows, n_out, mean + r0 * n_out,
                                                            spread ? spread + r0 : nullptr);
            check(hipGetLastError(), "ensemble_reduce launch");
        }
        check(hipStreamSynchronize(st), "ensemble predict (no allocation)");
    });
}

mbg_status mbg_mlp_predict(mbg_mlp* m, const float* x, size_t n, float* y, double* ms) {
    if (!m || !x || !y) return 3;
    if (n == 0) {
        if (ms) *ms = 0.0;
        return 0;
    }
    return mbg_mlp_ensemble_predict(&m, 1, x, n, y, nullptr, ms);
}

}  // extern "C"
