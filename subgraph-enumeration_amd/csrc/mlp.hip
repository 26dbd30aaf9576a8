// Learned attribute-order cost model: a 64 -> 128 -> 128 -> 1 MLP (ReLU) whose
// two hidden layers run on gfx950 matrix cores (v_mfma_f32_16x16x16_bf16: BF16
// operands, FP32 accumulation); FP32 master weights are trained with Adam.
//
// This has no counterpart in the reference: it stands in for the order choice
// of plan/optimizer.cpp:441-577 (optCostOrder), which stays the fallback.
//
// Layouts (row-major): X[N][64] bf16, W1[64][128] bf16, W2[128][128] bf16,
// H1/H2[N][128] bf16 (post-ReLU activations kept for the backward pass).
// Flat FP32 parameter/gradient vector: W1 | b1 | W2 | b2 | w3 | b3.
//
// MFMA 16x16x16 BF16 operand layout (lane l of the 64-wide wave):
//   A (16 x 16, row i, col k): lane l holds A[l % 16][4*(l/16) + q], q = 0..3
//   B (16 x 16, row k, col j): lane l holds B[4*(l/16) + q][l % 16]
//   D (16 x 16):               lane l holds D[4*(l/16) + q][l % 16]
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

#define HIP_TRY(x)                                                                         \
    do {                                                                                   \
        hipError_t _e = (x);                                                               \
        if (_e != hipSuccess) {                                                            \
            sge::set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #x); \
            return SGE_EDEVICE;                                                            \
        }                                                                                  \
    } while (0)

namespace {

constexpr int IN = 64, HID = 128;
constexpr int OFF_W1 = 0, OFF_B1 = IN * HID, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + HID * HID,
              OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + HID, NPARAM = OFF_B3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

__device__ inline floatx4 mfma16(bf16x4 a, bf16x4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// fp32 [n] -> bf16 [n]
__global__ void to_bf16_kernel(const float *in, uint16_t *out, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// One wave per 16 rows.  Writes y[N], H1, H2 (bf16).
__global__ void __launch_bounds__(256) forward_kernel(const uint16_t *X, uint32_t N, const uint16_t *W1,
                                                      const uint16_t *W2, const float *P, uint16_t *H1,
                                                      uint16_t *H2, float *y) {
    __shared__ uint16_t h1s[4][16][HID];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t r0 = (blockIdx.x * 4 + wv) * 16;
    if (r0 >= N) return;
    const int li = lane & 15, lq = lane >> 4;
    const float *b1 = P + OFF_B1, *b2 = P + OFF_B2, *w3 = P + OFF_W3;
    // ---- layer 1: z1 = X W1 + b1 ; h1 = relu(z1)
    for (int j = 0; j < HID / 16; ++j) {
        floatx4 acc = {0, 0, 0, 0};
        for (int kk = 0; kk < IN / 16; ++kk) {
            bf16x4 a, b;
            uint32_t row = r0 + li;
            for (int q = 0; q < 4; ++q) {
                int k = kk * 16 + lq * 4 + q;
                a[q] = row < N ? (short)X[(size_t)row * IN + k] : (short)0;
                b[q] = (short)W1[k * HID + j * 16 + li];
            }
            acc = mfma16(a, b, acc);
        }
        for (int q = 0; q < 4; ++q) {
            int r = lq * 4 + q, c = j * 16 + li;
            float z = acc[q] + b1[c];
            uint16_t h = f2bf(z > 0.f ? z : 0.f);
            h1s[wv][r][c] = h;
            if (r0 + r < N) H1[(size_t)(r0 + r) * HID + c] = h;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // ---- layer 2: z2 = h1 W2 + b2 ; h2 = relu(z2) ; layer 3 partial sums
    float part[4] = {0, 0, 0, 0};
    for (int j = 0; j < HID / 16; ++j) {
        floatx4 acc = {0, 0, 0, 0};
        for (int kk = 0; kk < HID / 16; ++kk) {
            bf16x4 a, b;
            for (int q = 0; q < 4; ++q) {
                int k = kk * 16 + lq * 4 + q;
                a[q] = (short)h1s[wv][li][k];
                b[q] = (short)W2[k * HID + j * 16 + li];
            }
            acc = mfma16(a, b, acc);
        }
        for (int q = 0; q < 4; ++q) {
            int r = lq * 4 + q, c = j * 16 + li;
            float z = acc[q] + b2[c];
            float h = z > 0.f ? z : 0.f;
            uint16_t hb = f2bf(h);
            if (r0 + r < N) H2[(size_t)(r0 + r) * HID + c] = hb;
            part[q] += bf2f(hb) * w3[c];
        }
    }
    // reduce over the 16 lanes that share lq (columns)
    for (int q = 0; q < 4; ++q)
        for (int o = 1; o < 16; o <<= 1) part[q] += __shfl_xor(part[q], o);
    if (li == 0)
        for (int q = 0; q < 4; ++q) {
            uint32_t r = r0 + lq * 4 + q;
            if (r < N) y[r] = part[q] + P[OFF_B3];
        }
}

// Backward, per 16-row tile: dy = 2 (y - t) / N (MSE); dz2 = dy w3 * [h2 > 0];
// dh1 = dz2 W2^T ; dz1 = dh1 * [h1 > 0].  Bias / output-layer gradients are
// accumulated with fp32 atomics; dz2 / dz1 are written as bf16 operands.
__global__ void __launch_bounds__(256) backward_rows_kernel(const float *y, const float *t, uint32_t N,
                                                            const uint16_t *W2, const float *P, const uint16_t *H1,
                                                            const uint16_t *H2, uint16_t *DZ2, uint16_t *DZ1,
                                                            float *G, float *loss) {
    __shared__ uint16_t dz2s[4][16][HID];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const uint32_t r0 = (blockIdx.x * 4 + wv) * 16;
    if (r0 >= N) return;
    const int li = lane & 15, lq = lane >> 4;
    const float *w3 = P + OFF_W3;
    // dy for the 4 rows this lane touches as a D-layout row owner
    float dy[4];
    for (int q = 0; q < 4; ++q) {
        uint32_t r = r0 + lq * 4 + q;
        dy[q] = r < N ? 2.f * (y[r] - t[r]) / (float)N : 0.f;
    }
    if (li == 0) {
        float l = 0, db3 = 0;
        for (int q = 0; q < 4; ++q) {
            uint32_t r = r0 + lq * 4 + q;
            if (r < N) {
                float e = y[r] - t[r];
                l += e * e / (float)N;
                db3 += dy[q];
            }
        }
        atomicAdd(loss, l);
        atomicAdd(G + OFF_B3, db3);
    }
    // dz2 tile: lane covers column c = j*16 + li for rows lq*4+q (D layout)
    for (int j = 0; j < HID / 16; ++j) {
        int c = j * 16 + li;
        float dw3 = 0, db2 = 0;
        for (int q = 0; q < 4; ++q) {
            int r = lq * 4 + q;
            float h2 = (r0 + r < N) ? bf2f(H2[(size_t)(r0 + r) * HID + c]) : 0.f;
            float g = (h2 > 0.f) ? dy[q] * w3[c] : 0.f;
            dw3 += dy[q] * h2;
            db2 += g;
            uint16_t gb = f2bf(g);
            dz2s[wv][r][c] = gb;
            if (r0 + r < N) DZ2[(size_t)(r0 + r) * HID + c] = gb;
        }
        // sum the 4 lq groups (lanes li, li+16, li+32, li+48)
        dw3 += __shfl_xor(dw3, 16);
        dw3 += __shfl_xor(dw3, 32);
        db2 += __shfl_xor(db2, 16);
        db2 += __shfl_xor(db2, 32);
        if (lq == 0) {
            atomicAdd(G + OFF_W3 + c, dw3);
            atomicAdd(G + OFF_B2 + c, db2);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // dh1 = dz2 W2^T : B[k][j'] = W2[j'][k]
    for (int j = 0; j < HID / 16; ++j) {
        floatx4 acc = {0, 0, 0, 0};
        for (int kk = 0; kk < HID / 16; ++kk) {
            bf16x4 a, b;
            for (int q = 0; q < 4; ++q) {
                int k = kk * 16 + lq * 4 + q;
                a[q] = (short)dz2s[wv][li][k];
                b[q] = (short)W2[(j * 16 + li) * HID + k];
            }
            acc = mfma16(a, b, acc);
        }
        int c = j * 16 + li;
        float db1 = 0;
        for (int q = 0; q < 4; ++q) {
            int r = lq * 4 + q;
            float h1 = (r0 + r < N) ? bf2f(H1[(size_t)(r0 + r) * HID + c]) : 0.f;
            float g = h1 > 0.f ? acc[q] : 0.f;
            db1 += g;
            if (r0 + r < N) DZ1[(size_t)(r0 + r) * HID + c] = f2bf(g);
        }
        db1 += __shfl_xor(db1, 16);
        db1 += __shfl_xor(db1, 32);
        if (lq == 0) atomicAdd(G + OFF_B1 + c, db1);
    }
}

// Weight gradients dW = A^T B summed over the batch: A[N][M] (M = 64 or 128),
// B[N][128]; one wave per 16x16 output tile; the batch is the MFMA K dimension.
__global__ void __launch_bounds__(64) weight_grad_kernel(const uint16_t *A, uint32_t M, const uint16_t *B, uint32_t N,
                                                         float *dW) {
    const int lane = threadIdx.x & 63;
    const int li = lane & 15, lq = lane >> 4;
    const uint32_t tiles_n = HID / 16;
    const uint32_t ti = blockIdx.x / tiles_n, tj = blockIdx.x % tiles_n;
    if (ti * 16 >= M) return;
    floatx4 acc = {0, 0, 0, 0};
    for (uint32_t n0 = 0; n0 < N; n0 += 16) {
        bf16x4 a, b;
        for (int q = 0; q < 4; ++q) {
            uint32_t n = n0 + lq * 4 + q;
            a[q] = n < N ? (short)A[(size_t)n * M + ti * 16 + li] : (short)0;
            b[q] = n < N ? (short)B[(size_t)n * HID + tj * 16 + li] : (short)0;
        }
        acc = mfma16(a, b, acc);
    }
    for (int q = 0; q < 4; ++q) dW[(ti * 16 + lq * 4 + q) * HID + tj * 16 + li] = acc[q];
}

__global__ void zero_kernel(float *p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// Adam on the fp32 master vector, then refresh the bf16 copies of W1, W2.
__global__ void adam_kernel(float *P, const float *G, float *m, float *v, uint32_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2, uint16_t *W1, uint16_t *W2) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float g = G[i];
    float mi = b1 * m[i] + (1.f - b1) * g;
    float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    float p = P[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    P[i] = p;
    if (i < (uint32_t)OFF_B1) W1[i - OFF_W1] = f2bf(p);
    else if (i >= (uint32_t)OFF_W2 && i < (uint32_t)OFF_B2) W2[i - OFF_W2] = f2bf(p);
}

}  // namespace

struct sge_mlp {
    float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *W1 = nullptr, *W2 = nullptr;
    // per-batch buffers
    uint32_t cap = 0;
    float *Xf = nullptr, *y = nullptr, *t = nullptr, *loss = nullptr;
    uint16_t *X = nullptr, *H1 = nullptr, *H2 = nullptr, *DZ2 = nullptr, *DZ1 = nullptr;
    hipStream_t stream = nullptr;
    uint32_t step = 0;
};

// Rows per call: keeps N * IN and every grid size below 2^31.
constexpr uint32_t MLP_MAX_ROWS = 1u << 24;

static int mlp_reserve(sge_mlp *M, uint32_t N) {
    if (N > MLP_MAX_ROWS) {
        sge::set_error("MLP: at most 2^24 rows per call");
        return SGE_ERANGE;
    }
    if (N <= M->cap) return SGE_OK;
    uint32_t c = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(N, 2ull * M->cap), MLP_MAX_ROWS);
    c = (c + 63) / 64 * 64;
    // released first, so that a failed allocation below leaves no freed pointer behind
    // (cap 0: the next call allocates again; destroy frees what was allocated)
    auto drop = [](auto *&q) {
        (void)hipFree(q);
        q = nullptr;
    };
    drop(M->Xf), drop(M->y), drop(M->t), drop(M->X), drop(M->H1), drop(M->H2), drop(M->DZ2), drop(M->DZ1);
    M->cap = 0;
    HIP_TRY(hipMalloc(&M->Xf, sizeof(float) * c * IN));
    HIP_TRY(hipMalloc(&M->y, sizeof(float) * c));
    HIP_TRY(hipMalloc(&M->t, sizeof(float) * c));
    HIP_TRY(hipMalloc(&M->X, sizeof(uint16_t) * c * IN));
    HIP_TRY(hipMalloc(&M->H1, sizeof(uint16_t) * c * HID));
    HIP_TRY(hipMalloc(&M->H2, sizeof(uint16_t) * c * HID));
    HIP_TRY(hipMalloc(&M->DZ2, sizeof(uint16_t) * c * HID));
    HIP_TRY(hipMalloc(&M->DZ1, sizeof(uint16_t) * c * HID));
    M->cap = c;
    return SGE_OK;
}

static int refresh_bf16(sge_mlp *M) {
    to_bf16_kernel<<<(IN * HID + 255) / 256, 256, 0, M->stream>>>(M->P + OFF_W1, M->W1, IN * HID);
    to_bf16_kernel<<<(HID * HID + 255) / 256, 256, 0, M->stream>>>(M->P + OFF_W2, M->W2, HID * HID);
    HIP_TRY(hipGetLastError());
    return SGE_OK;
}

extern "C" int sge_mlp_num_params(void) { return NPARAM; }
extern "C" int sge_mlp_input_dim(void) { return IN; }

extern "C" void sge_mlp_destroy(sge_mlp *M) {
    if (!M) return;
    (void)hipFree(M->P); (void)hipFree(M->G); (void)hipFree(M->m); (void)hipFree(M->v); (void)hipFree(M->W1); (void)hipFree(M->W2);
    (void)hipFree(M->Xf); (void)hipFree(M->y); (void)hipFree(M->t); (void)hipFree(M->loss); (void)hipFree(M->X); (void)hipFree(M->H1);
    (void)hipFree(M->H2); (void)hipFree(M->DZ2); (void)hipFree(M->DZ1);
    if (M->stream) (void)hipStreamDestroy(M->stream);
    delete M;
}

extern "C" int sge_mlp_create(uint64_t seed, sge_mlp **out) {
    SGE_API_BEGIN
    if (!out) {
        sge::set_error("sge_mlp_create: null output");
        return SGE_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        sge::set_error("no HIP device available for the MLP");
        return SGE_EDEVICE;
    }
    sge_mlp *M = new sge_mlp();
    auto fail = [&](const char *w) {
        sge::set_error(std::string("HIP allocation failed: ") + w);
        sge_mlp_destroy(M);
        return SGE_EDEVICE;
    };
    if (hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) != hipSuccess) return fail("stream");
    if (hipMalloc(&M->P, sizeof(float) * NPARAM) || hipMalloc(&M->G, sizeof(float) * NPARAM) ||
        hipMalloc(&M->m, sizeof(float) * NPARAM) || hipMalloc(&M->v, sizeof(float) * NPARAM) ||
        hipMalloc(&M->W1, sizeof(uint16_t) * IN * HID) || hipMalloc(&M->W2, sizeof(uint16_t) * HID * HID) ||
        hipMalloc(&M->loss, sizeof(float)))
        return fail("parameters");
    // He-uniform init from a seeded host RNG
    std::vector<float> p(NPARAM, 0.f);
    sge::Rng rng(seed);
    auto fill = [&](int off, int n, int fan_in) {
        float a = std::sqrt(6.0f / fan_in);
        for (int i = 0; i < n; ++i) p[off + i] = (float)((2.0 * rng.uniform() - 1.0) * a);
    };
    fill(OFF_W1, IN * HID, IN);
    fill(OFF_W2, HID * HID, HID);
    fill(OFF_W3, HID, HID);
    if (hipMemcpy(M->P, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) ||
        hipMemset(M->m, 0, sizeof(float) * NPARAM) || hipMemset(M->v, 0, sizeof(float) * NPARAM))
        return fail("init copy");
    if (refresh_bf16(M)) return fail("bf16 copy");
    if (hipStreamSynchronize(M->stream)) return fail("sync");
    *out = M;
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_mlp_set_params(sge_mlp *M, const float *params, size_t n) {
    SGE_API_BEGIN
    if (!M || !params || n != (size_t)NPARAM) {
        sge::set_error("sge_mlp_set_params: bad arguments");
        return SGE_EINVAL;
    }
    HIP_TRY(hipMemcpy(M->P, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
    int rc = refresh_bf16(M);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(M->stream));
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_mlp_get_params(sge_mlp *M, float *params, size_t n) {
    SGE_API_BEGIN
    if (!M || !params || n != (size_t)NPARAM) {
        sge::set_error("sge_mlp_get_params: bad arguments");
        return SGE_EINVAL;
    }
    HIP_TRY(hipMemcpy(params, M->P, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
    return SGE_OK;
    SGE_API_END
}

static int upload_x(sge_mlp *M, const float *X, uint32_t N) {
    int rc = mlp_reserve(M, N);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(M->Xf, X, sizeof(float) * N * IN, hipMemcpyHostToDevice, M->stream));
    to_bf16_kernel<<<(N * IN + 255) / 256, 256, 0, M->stream>>>(M->Xf, M->X, (size_t)N * IN);
    HIP_TRY(hipGetLastError());
    return SGE_OK;
}

static int run_forward(sge_mlp *M, uint32_t N) {
    uint32_t tiles = (N + 15) / 16;
    forward_kernel<<<(tiles + 3) / 4, 256, 0, M->stream>>>(M->X, N, M->W1, M->W2, M->P, M->H1, M->H2, M->y);
    HIP_TRY(hipGetLastError());
    return SGE_OK;
}

extern "C" int sge_mlp_forward(sge_mlp *M, const float *X, uint32_t N, float *y) {
    SGE_API_BEGIN
    if (!M || !X || !y || N == 0) {
        sge::set_error("sge_mlp_forward: bad arguments");
        return SGE_EINVAL;
    }
    int rc = upload_x(M, X, N);
    if (rc) return rc;
    rc = run_forward(M, N);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(y, M->y, sizeof(float) * N, hipMemcpyDeviceToHost, M->stream));
    HIP_TRY(hipStreamSynchronize(M->stream));
    return SGE_OK;
    SGE_API_END
}

// forward + backward into M->G (no update)
static int run_grads(sge_mlp *M, const float *X, const float *t, uint32_t N) {
    int rc = upload_x(M, X, N);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(M->t, t, sizeof(float) * N, hipMemcpyHostToDevice, M->stream));
    rc = run_forward(M, N);
    if (rc) return rc;
    zero_kernel<<<(NPARAM + 255) / 256, 256, 0, M->stream>>>(M->G, NPARAM);
    zero_kernel<<<1, 64, 0, M->stream>>>(M->loss, 1);
    uint32_t tiles = (N + 15) / 16;
    backward_rows_kernel<<<(tiles + 3) / 4, 256, 0, M->stream>>>(M->y, M->t, N, M->W2, M->P, M->H1, M->H2, M->DZ2,
                                                                 M->DZ1, M->G, M->loss);
    weight_grad_kernel<<<(HID / 16) * (HID / 16), 64, 0, M->stream>>>(M->H1, HID, M->DZ2, N, M->G + OFF_W2);
    weight_grad_kernel<<<(IN / 16) * (HID / 16), 64, 0, M->stream>>>(M->X, IN, M->DZ1, N, M->G + OFF_W1);
    HIP_TRY(hipGetLastError());
    return SGE_OK;
}

extern "C" int sge_mlp_gradients(sge_mlp *M, const float *X, const float *t, uint32_t N, float *grads, float *loss) {
    SGE_API_BEGIN
    if (!M || !X || !t || !grads || N == 0) {
        sge::set_error("sge_mlp_gradients: bad arguments");
        return SGE_EINVAL;
    }
    int rc = run_grads(M, X, t, N);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(grads, M->G, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, M->stream));
    if (loss) HIP_TRY(hipMemcpyAsync(loss, M->loss, sizeof(float), hipMemcpyDeviceToHost, M->stream));
    HIP_TRY(hipStreamSynchronize(M->stream));
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_mlp_train_step(sge_mlp *M, const float *X, const float *t, uint32_t N, float lr, float *loss) {
    SGE_API_BEGIN
    if (!M || !X || !t || N == 0) {
        sge::set_error("sge_mlp_train_step: bad arguments");
        return SGE_EINVAL;
    }
    int rc = run_grads(M, X, t, N);
    if (rc) return rc;
    M->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float c1 = 1.f - std::pow(b1, (float)M->step), c2 = 1.f - std::pow(b2, (float)M->step);
    adam_kernel<<<(NPARAM + 255) / 256, 256, 0, M->stream>>>(M->P, M->G, M->m, M->v, NPARAM, lr, b1, b2, eps, c1, c2,
                                                              M->W1, M->W2);
    HIP_TRY(hipGetLastError());
    if (loss) HIP_TRY(hipMemcpyAsync(loss, M->loss, sizeof(float), hipMemcpyDeviceToHost, M->stream));
    HIP_TRY(hipStreamSynchronize(M->stream));
    return SGE_OK;
    SGE_API_END
}
