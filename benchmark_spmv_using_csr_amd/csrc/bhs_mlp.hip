// bhs_mlp.hip -- forward, backward and Adam kernels of the tuner's MLP, the
// fused on-device inference kernel, and their C ABI.
//
// Every reduction runs in a fixed order inside one thread or one block (no
// atomics), so training is bitwise reproducible.  The kernels are templated on
// the value type: training runs in float, and a double instantiation backs the
// gradient check against the CPU FP64 oracle.
#include "bhs_device.h"
#include "bhs_tuner.h"
#include "../../include/bhsparse_spmv.h"

#include <cstring>
#include <new>
#include <string>
#include <vector>

namespace bhs {

int mlp_set_error(int code, const std::string &msg);   // defined in bhs_api.hip

// Y[n, o] = act(sum_i X[n, i] W[o, i] + b[o]); optional input standardisation
// (xs = (x - mean[i]) / std[i]) for the first layer.
template <typename T>
__global__ __launch_bounds__(256) void mlp_linear_fwd(const T *__restrict__ X, int N, int I,
                                                     const T *__restrict__ W,
                                                     const T *__restrict__ b, int O, int relu,
                                                     const T *__restrict__ norm,
                                                     T *__restrict__ Y)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * O) return;
    const int n = (int)(t / O), o = (int)(t % O);
    T acc = b[o];
    for (int i = 0; i < I; ++i) {
        T xv = X[(long long)n * I + i];
        if (norm) xv = (xv - norm[i]) / norm[I + i];
        acc += xv * W[(long long)o * I + i];
    }
    Y[t] = (relu && acc < T(0)) ? T(0) : acc;
}

// Masked mean squared error and its gradient dOut = 2 mask (out - target) / sum(mask).
// One block; fixed-order partial sums, then a fixed tree over the block.
template <typename T>
__global__ __launch_bounds__(256) void mlp_loss_grad(const T *__restrict__ out,
                                                    const T *__restrict__ target,
                                                    const T *__restrict__ mask, long long NK,
                                                    T *__restrict__ dout, double *__restrict__ loss)
{
    __shared__ double s_l[256], s_m[256];
    double l = 0, mk = 0;
    for (long long i = threadIdx.x; i < NK; i += blockDim.x) {
        const double d = (double)out[i] - (double)target[i];
        l += (double)mask[i] * d * d;
        mk += (double)mask[i];
    }
    s_l[threadIdx.x] = l;
    s_m[threadIdx.x] = mk;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s_l[threadIdx.x] += s_l[threadIdx.x + s];
            s_m[threadIdx.x] += s_m[threadIdx.x + s];
        }
        __syncthreads();
    }
    const double msum = s_m[0] > 0 ? s_m[0] : 1.0;
    for (long long i = threadIdx.x; i < NK; i += blockDim.x)
        dout[i] = (T)(2.0 * (double)mask[i] * ((double)out[i] - (double)target[i]) / msum);
    if (threadIdx.x == 0) *loss = s_l[0] / msum;
}

// dX[n, i] = (sum_o dY[n, o] W[o, i]) * relu'(Xact[n, i]); Xact is the
// activation the layer consumed (its derivative is 1 where it is > 0).
template <typename T>
__global__ __launch_bounds__(256) void mlp_linear_bwd_x(const T *__restrict__ dY, int N, int O,
                                                       const T *__restrict__ W, int I,
                                                       const T *__restrict__ Xact,
                                                       T *__restrict__ dX)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)N * I) return;
    const int n = (int)(t / I), i = (int)(t % I);
    T acc = T(0);
    for (int o = 0; o < O; ++o) acc += dY[(long long)n * O + o] * W[(long long)o * I + i];
    dX[t] = Xact[t] > T(0) ? acc : T(0);
}

// dW[o, i] = sum_n dY[n, o] Xs[n, i] (Xs standardised when norm != null),
// db[o] = sum_n dY[n, o]; thread (o, i) sums over n in order, i == I gives db.
template <typename T>
__global__ __launch_bounds__(256) void mlp_linear_bwd_w(const T *__restrict__ dY, int N, int O,
                                                       const T *__restrict__ X, int I,
                                                       const T *__restrict__ norm,
                                                       T *__restrict__ dW, T *__restrict__ db)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= O * (I + 1)) return;
    const int o = t / (I + 1), i = t % (I + 1);
    T acc = T(0);
    if (i == I) {
        for (int n = 0; n < N; ++n) acc += dY[(long long)n * O + o];
        db[o] = acc;
    } else {
        const T mu = norm ? norm[i] : T(0), sd = norm ? norm[I + i] : T(1);
        for (int n = 0; n < N; ++n) {
            T xv = X[(long long)n * I + i];
            if (norm) xv = (xv - mu) / sd;
            acc += dY[(long long)n * O + o] * xv;
        }
        dW[(long long)o * I + i] = acc;
    }
}

// Adam with decoupled weight decay, step t >= 1.
__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ p, const float *__restrict__ g,
                                                  float *__restrict__ m1, float *__restrict__ m2,
                                                  long long n, float lr, float b1, float b2,
                                                  float eps, float wd, int t)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float ah = a / (1.f - powf(b1, (float)t));
    const float bh = b / (1.f - powf(b2, (float)t));
    p[i] = p[i] - lr * (ah / (sqrtf(bh) + eps) + wd * p[i]);
}

// One matrix per block of 64 threads: integer row statistics -> raw features
// -> transform -> standardise -> MLP -> argmin over the classes with valid[k].
__global__ __launch_bounds__(64) void tuner_infer_kernel(
    const unsigned long long *__restrict__ stats, const int *__restrict__ mn,
    const int *__restrict__ is_double, int n_mat, const float *__restrict__ params,
    const float *__restrict__ norm, int h, int K, const int *__restrict__ valid,
    int *__restrict__ choice, float *__restrict__ scores)
{
    const int mat = blockIdx.x;
    if (mat >= n_mat) return;
    __shared__ float s_x[kTunerFeatures], s_h1[kMaxHidden], s_h2[kMaxHidden], s_o[kMaxClasses];
    const int tid = threadIdx.x;
    if (tid == 0) {
        double raw[kTunerFeatures];
        tuner_raw(stats + (long long)mat * 8, mn[2 * mat], mn[2 * mat + 1], is_double[mat], raw);
        float f[kTunerFeatures];
        tuner_transform(raw, f);
        for (int i = 0; i < kTunerFeatures; ++i) s_x[i] = (f[i] - norm[i]) / norm[kTunerFeatures + i];
    }
    __syncthreads();
    const int I = kTunerFeatures;
    const float *W1 = params, *b1 = W1 + h * I, *W2 = b1 + h, *b2 = W2 + h * h, *W3 = b2 + h,
                *b3 = W3 + K * h;
    for (int o = tid; o < h; o += blockDim.x) {
        float a = b1[o];
        for (int i = 0; i < I; ++i) a += s_x[i] * W1[o * I + i];
        s_h1[o] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int o = tid; o < h; o += blockDim.x) {
        float a = b2[o];
        for (int i = 0; i < h; ++i) a += s_h1[i] * W2[o * h + i];
        s_h2[o] = a > 0.f ? a : 0.f;
    }
    __syncthreads();
    for (int o = tid; o < K; o += blockDim.x) {
        float a = b3[o];
        for (int i = 0; i < h; ++i) a += s_h2[i] * W3[o * h + i];
        s_o[o] = a;
        if (scores) scores[(long long)mat * K + o] = a;
    }
    __syncthreads();
    if (tid == 0) {
        int best = -1;
        float bv = 0.f;
        for (int k = 0; k < K; ++k)
            if (valid[k] && (best < 0 || s_o[k] < bv)) { best = k; bv = s_o[k]; }
        choice[mat] = best;
    }
}

hipError_t launch_tuner_infer(const unsigned long long *stats, const int *mn, const int *is_double,
                              int n_mat, const Mlp &net, int K, const int *valid, int *choice,
                              float *scores, hipStream_t stream)
{
    hipLaunchKernelGGL(tuner_infer_kernel, dim3(n_mat), dim3(64), 0, stream, stats, mn, is_double,
                       n_mat, net.d_params, net.d_norm, net.h, K, valid, choice, scores);
    return hipGetLastError();
}

namespace {

unsigned grid_for(long long n) { return (unsigned)((n + 255) / 256); }

template <typename T>
struct Work {
    T *X = nullptr, *Tg = nullptr, *Mk = nullptr, *H1 = nullptr, *H2 = nullptr, *O = nullptr;
    T *dO = nullptr, *dH2 = nullptr, *dH1 = nullptr, *grad = nullptr, *params = nullptr,
      *norm = nullptr;
    double *loss = nullptr;
    std::vector<void *> owned;
    ~Work()
    {
        for (void *p : owned) (void)hipFree(p);
    }
    template <typename U>
    hipError_t alloc(U **p, long long n)
    {
        hipError_t e = hipMalloc((void **)p, sizeof(U) * (size_t)(n > 0 ? n : 1));
        if (e == hipSuccess) owned.push_back(*p);
        return e;
    }
};

// Forward + loss + backward for N rows on `stream`; params/norm/X/T/M already
// on the device in w.  Writes w.grad and *w.loss.
template <typename T>
void mlp_fwd_bwd(Work<T> &w, int N, int I, int h, int K, hipStream_t st)
{
    T *W1 = w.params, *b1 = W1 + (long long)h * I, *W2 = b1 + h, *b2 = W2 + (long long)h * h,
      *W3 = b2 + h, *b3 = W3 + (long long)K * h;
    T *gW1 = w.grad, *gb1 = gW1 + (long long)h * I, *gW2 = gb1 + h, *gb2 = gW2 + (long long)h * h,
      *gW3 = gb2 + h, *gb3 = gW3 + (long long)K * h;
    hipLaunchKernelGGL(mlp_linear_fwd<T>, dim3(grid_for((long long)N * h)), dim3(256), 0, st, w.X,
                       N, I, W1, b1, h, 1, w.norm, w.H1);
    hipLaunchKernelGGL(mlp_linear_fwd<T>, dim3(grid_for((long long)N * h)), dim3(256), 0, st, w.H1,
                       N, h, W2, b2, h, 1, (const T *)nullptr, w.H2);
    hipLaunchKernelGGL(mlp_linear_fwd<T>, dim3(grid_for((long long)N * K)), dim3(256), 0, st, w.H2,
                       N, h, W3, b3, K, 0, (const T *)nullptr, w.O);
    hipLaunchKernelGGL(mlp_loss_grad<T>, dim3(1), dim3(256), 0, st, w.O, w.Tg, w.Mk,
                       (long long)N * K, w.dO, w.loss);
    hipLaunchKernelGGL(mlp_linear_bwd_w<T>, dim3(grid_for((long long)K * (h + 1))), dim3(256), 0,
                       st, w.dO, N, K, w.H2, h, (const T *)nullptr, gW3, gb3);
    hipLaunchKernelGGL(mlp_linear_bwd_x<T>, dim3(grid_for((long long)N * h)), dim3(256), 0, st,
                       w.dO, N, K, W3, h, w.H2, w.dH2);
    hipLaunchKernelGGL(mlp_linear_bwd_w<T>, dim3(grid_for((long long)h * (h + 1))), dim3(256), 0,
                       st, w.dH2, N, h, w.H1, h, (const T *)nullptr, gW2, gb2);
    hipLaunchKernelGGL(mlp_linear_bwd_x<T>, dim3(grid_for((long long)N * h)), dim3(256), 0, st,
                       w.dH2, N, h, W2, h, w.H1, w.dH1);
    hipLaunchKernelGGL(mlp_linear_bwd_w<T>, dim3(grid_for((long long)h * (I + 1))), dim3(256), 0,
                       st, w.dH1, N, h, w.X, I, w.norm, gW1, gb1);
}

#define MLP_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return mlp_set_error(e_ == hipErrorOutOfMemory ? BHSPARSE_OUT_OF_MEMORY          \
                                                           : BHSPARSE_HIP_ERROR,            \
                                 std::string(#call) + ": " + hipGetErrorString(e_));        \
    } while (0)

template <typename T>
int setup_work(Work<T> &w, int N, int I, int h, int K, const T *x, const T *tg, const T *mk,
               hipStream_t st)
{
    const long long P = mlp_param_count(I, h, K);
    MLP_HIP(w.alloc(&w.X, (long long)N * I));
    MLP_HIP(w.alloc(&w.Tg, (long long)N * K));
    MLP_HIP(w.alloc(&w.Mk, (long long)N * K));
    MLP_HIP(w.alloc(&w.H1, (long long)N * h));
    MLP_HIP(w.alloc(&w.H2, (long long)N * h));
    MLP_HIP(w.alloc(&w.O, (long long)N * K));
    MLP_HIP(w.alloc(&w.dO, (long long)N * K));
    MLP_HIP(w.alloc(&w.dH2, (long long)N * h));
    MLP_HIP(w.alloc(&w.dH1, (long long)N * h));
    MLP_HIP(w.alloc(&w.grad, P));
    MLP_HIP(w.alloc(&w.loss, 1));
    MLP_HIP(hipMemcpyAsync(w.X, x, sizeof(T) * (size_t)N * I, hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(w.Tg, tg, sizeof(T) * (size_t)N * K, hipMemcpyHostToDevice, st));
    MLP_HIP(hipMemcpyAsync(w.Mk, mk, sizeof(T) * (size_t)N * K, hipMemcpyHostToDevice, st));
    return BHSPARSE_SUCCESS;
}

int check_net(bhs_mlp_t net, Mlp **out)
{
    if (!net) return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "null network");
    *out = reinterpret_cast<Mlp *>(net);
    MLP_HIP(hipSetDevice((*out)->device));
    return BHSPARSE_SUCCESS;
}

} // namespace
} // namespace bhs

using namespace bhs;

extern "C" {

int bhs_tuner_transform(const double *raw8, float *out8)
{
    if (!raw8 || !out8) return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "null features");
    tuner_transform(raw8, out8);
    return BHSPARSE_SUCCESS;
}

int bhs_mlp_create(bhs_mlp_t *net, int device, int n_in, int n_hidden, int n_out)
try {
    if (!net || n_in != kTunerFeatures || n_hidden < 1 || n_hidden > kMaxHidden || n_out < 1 ||
        n_out > kMaxClasses)
        return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "network shape out of range");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return mlp_set_error(BHSPARSE_UNSUPPORTED_DEVICE, "no such device");
    MLP_HIP(hipSetDevice(device));
    Mlp *m = new (std::nothrow) Mlp;
    if (!m) return mlp_set_error(BHSPARSE_OUT_OF_MEMORY, "host allocation failed");
    m->device = device;
    m->n_in = n_in;
    m->h = n_hidden;
    m->n_out = n_out;
    m->n_params = mlp_param_count(n_in, n_hidden, n_out);
    hipError_t e = hipMalloc(&m->d_params, sizeof(float) * (size_t)m->n_params);
    if (e == hipSuccess) e = hipMalloc(&m->d_norm, sizeof(float) * 2 * (size_t)n_in);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemset(m->d_params, 0, sizeof(float) * (size_t)m->n_params);
    if (e == hipSuccess) {
        std::vector<float> nm(2 * n_in, 0.f);
        for (int i = 0; i < n_in; ++i) nm[n_in + i] = 1.f;
        e = hipMemcpy(m->d_norm, nm.data(), sizeof(float) * nm.size(), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        if (m->d_params) (void)hipFree(m->d_params);
        if (m->d_norm) (void)hipFree(m->d_norm);
        if (m->stream) (void)hipStreamDestroy(m->stream);
        delete m;
        return mlp_set_error(BHSPARSE_HIP_ERROR, std::string("mlp create: ") + hipGetErrorString(e));
    }
    *net = reinterpret_cast<bhs_mlp_t>(m);
    return BHSPARSE_SUCCESS;
}
catch (const std::bad_alloc &) {   // no exception crosses the C boundary
    return mlp_set_error(BHSPARSE_OUT_OF_MEMORY, "host allocation failed");
} catch (...) {
    return mlp_set_error(BHSPARSE_HIP_ERROR, "internal error");
}

int bhs_mlp_destroy(bhs_mlp_t net)
{
    Mlp *m;
    if (int rc = check_net(net, &m)) return rc;
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(m->d_params);
    (void)hipFree(m->d_norm);
    (void)hipStreamDestroy(m->stream);
    delete m;
    return BHSPARSE_SUCCESS;
}

long long bhs_mlp_param_count(int n_in, int n_hidden, int n_out)
{
    return mlp_param_count(n_in, n_hidden, n_out);
}

int bhs_mlp_set_params(bhs_mlp_t net, const float *params, long long n, const float *norm_mean,
                       const float *norm_std)
try {
    Mlp *m;
    if (int rc = check_net(net, &m)) return rc;
    if (!params || n != m->n_params || !norm_mean || !norm_std)
        return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "parameter count mismatch");
    MLP_HIP(hipMemcpy(m->d_params, params, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    std::vector<float> nm(2 * m->n_in);
    for (int i = 0; i < m->n_in; ++i) {
        if (!(norm_std[i] > 0)) return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "std must be > 0");
        nm[i] = norm_mean[i];
        nm[m->n_in + i] = norm_std[i];
    }
    MLP_HIP(hipMemcpy(m->d_norm, nm.data(), sizeof(float) * nm.size(), hipMemcpyHostToDevice));
    return BHSPARSE_SUCCESS;
}
catch (const std::bad_alloc &) {   // no exception crosses the C boundary
    return mlp_set_error(BHSPARSE_OUT_OF_MEMORY, "host allocation failed");
} catch (...) {
    return mlp_set_error(BHSPARSE_HIP_ERROR, "internal error");
}

int bhs_mlp_get_params(bhs_mlp_t net, float *params, long long n)
{
    Mlp *m;
    if (int rc = check_net(net, &m)) return rc;
    if (!params || n != m->n_params)
        return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "parameter count mismatch");
    MLP_HIP(hipStreamSynchronize(m->stream));
    MLP_HIP(hipMemcpy(params, m->d_params, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    return BHSPARSE_SUCCESS;
}

int bhs_mlp_forward(bhs_mlp_t net, const float *x, int n_rows, float *out)
try {
    Mlp *m;
    if (int rc = check_net(net, &m)) return rc;
    if (!x || !out || n_rows < 1) return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "bad batch");
    Work<float> w;
    std::vector<float> zeros((size_t)n_rows * m->n_out, 0.f);
    if (int rc = setup_work<float>(w, n_rows, m->n_in, m->h, m->n_out, x, zeros.data(),
                                   zeros.data(), m->stream))
        return rc;
    const int N = n_rows, I = m->n_in, h = m->h, K = m->n_out;
    const float *W1 = m->d_params, *b1 = W1 + h * I, *W2 = b1 + h, *b2 = W2 + h * h, *W3 = b2 + h,
                *b3 = W3 + K * h;
    hipLaunchKernelGGL(mlp_linear_fwd<float>, dim3(grid_for((long long)N * h)), dim3(256), 0,
                       m->stream, w.X, N, I, W1, b1, h, 1, (const float *)m->d_norm, w.H1);
    hipLaunchKernelGGL(mlp_linear_fwd<float>, dim3(grid_for((long long)N * h)), dim3(256), 0,
                       m->stream, w.H1, N, h, W2, b2, h, 1, (const float *)nullptr, w.H2);
    hipLaunchKernelGGL(mlp_linear_fwd<float>, dim3(grid_for((long long)N * K)), dim3(256), 0,
                       m->stream, w.H2, N, h, W3, b3, K, 0, (const float *)nullptr, w.O);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(out, w.O, sizeof(float) * (size_t)N * K, hipMemcpyDeviceToHost,
                           m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return BHSPARSE_SUCCESS;
}
catch (const std::bad_alloc &) {   // no exception crosses the C boundary
    return mlp_set_error(BHSPARSE_OUT_OF_MEMORY, "host allocation failed");
} catch (...) {
    return mlp_set_error(BHSPARSE_HIP_ERROR, "internal error");
}

int bhs_mlp_loss_grad(bhs_mlp_t net, const float *x, const float *target, const float *mask,
                      int n_rows, double *loss, float *grad)
try {
    Mlp *m;
    if (int rc = check_net(net, &m)) return rc;
    if (!x || !target || !mask || !loss || !grad || n_rows < 1)
        return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "bad batch");
    Work<float> w;
    if (int rc = setup_work<float>(w, n_rows, m->n_in, m->h, m->n_out, x, target, mask, m->stream))
        return rc;
    w.params = m->d_params;
    w.norm = m->d_norm;
    mlp_fwd_bwd<float>(w, n_rows, m->n_in, m->h, m->n_out, m->stream);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(grad, w.grad, sizeof(float) * (size_t)m->n_params,
                           hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipMemcpyAsync(loss, w.loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return BHSPARSE_SUCCESS;
}
catch (const std::bad_alloc &) {   // no exception crosses the C boundary
    return mlp_set_error(BHSPARSE_OUT_OF_MEMORY, "host allocation failed");
} catch (...) {
    return mlp_set_error(BHSPARSE_HIP_ERROR, "internal error");
}

int bhs_mlp_train(bhs_mlp_t net, const float *x, const float *target, const float *mask,
                  int n_rows, int epochs, float lr, float weight_decay, double *loss_hist)
try {
    Mlp *m;
    if (int rc = check_net(net, &m)) return rc;
    if (!x || !target || !mask || n_rows < 1 || epochs < 1 || !(lr > 0))
        return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "bad training arguments");
    Work<float> w;
    if (int rc = setup_work<float>(w, n_rows, m->n_in, m->h, m->n_out, x, target, mask, m->stream))
        return rc;
    w.params = m->d_params;
    w.norm = m->d_norm;
    float *m1 = nullptr, *m2 = nullptr;
    double *hist = nullptr;
    MLP_HIP(w.alloc(&m1, m->n_params));
    MLP_HIP(w.alloc(&m2, m->n_params));
    MLP_HIP(w.alloc(&hist, epochs));
    MLP_HIP(hipMemsetAsync(m1, 0, sizeof(float) * (size_t)m->n_params, m->stream));
    MLP_HIP(hipMemsetAsync(m2, 0, sizeof(float) * (size_t)m->n_params, m->stream));
    for (int ep = 1; ep <= epochs; ++ep) {
        w.loss = hist + (ep - 1);
        mlp_fwd_bwd<float>(w, n_rows, m->n_in, m->h, m->n_out, m->stream);
        hipLaunchKernelGGL(adam_kernel, dim3(grid_for(m->n_params)), dim3(256), 0, m->stream,
                           m->d_params, w.grad, m1, m2, m->n_params, lr, 0.9f, 0.999f, 1e-8f,
                           weight_decay, ep);
    }
    MLP_HIP(hipGetLastError());
    if (loss_hist)
        MLP_HIP(hipMemcpyAsync(loss_hist, hist, sizeof(double) * (size_t)epochs,
                               hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return BHSPARSE_SUCCESS;
}
catch (const std::bad_alloc &) {   // no exception crosses the C boundary
    return mlp_set_error(BHSPARSE_OUT_OF_MEMORY, "host allocation failed");
} catch (...) {
    return mlp_set_error(BHSPARSE_HIP_ERROR, "internal error");
}

int bhs_mlp_loss_grad_f64(int device, int n_in, int n_hidden, int n_out, const double *params,
                          const double *norm, const double *x, const double *target,
                          const double *mask, int n_rows, double *loss, double *grad)
try {
    if (!params || !norm || !x || !target || !mask || !loss || !grad || n_rows < 1 ||
        n_in != kTunerFeatures || n_hidden < 1 || n_hidden > kMaxHidden || n_out < 1 ||
        n_out > kMaxClasses)
        return mlp_set_error(BHSPARSE_INVALID_ARGUMENT, "bad arguments");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return mlp_set_error(BHSPARSE_UNSUPPORTED_DEVICE, "no such device");
    MLP_HIP(hipSetDevice(device));
    hipStream_t st;
    MLP_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    int rc = BHSPARSE_SUCCESS;
    {
        Work<double> w;
        const long long P = mlp_param_count(n_in, n_hidden, n_out);
        rc = setup_work<double>(w, n_rows, n_in, n_hidden, n_out, x, target, mask, st);
        hipError_t e = hipSuccess;
        if (!rc) e = w.alloc(&w.params, P);
        if (!rc && e == hipSuccess) e = w.alloc(&w.norm, 2 * (long long)n_in);
        if (!rc && e == hipSuccess)
            e = hipMemcpyAsync(w.params, params, sizeof(double) * (size_t)P, hipMemcpyHostToDevice, st);
        if (!rc && e == hipSuccess)
            e = hipMemcpyAsync(w.norm, norm, sizeof(double) * 2 * (size_t)n_in, hipMemcpyHostToDevice, st);
        if (!rc && e == hipSuccess) {
            mlp_fwd_bwd<double>(w, n_rows, n_in, n_hidden, n_out, st);
            e = hipGetLastError();
        }
        if (!rc && e == hipSuccess)
            e = hipMemcpyAsync(grad, w.grad, sizeof(double) * (size_t)P, hipMemcpyDeviceToHost, st);
        if (!rc && e == hipSuccess)
            e = hipMemcpyAsync(loss, w.loss, sizeof(double), hipMemcpyDeviceToHost, st);
        if (!rc && e == hipSuccess) e = hipStreamSynchronize(st);
        if (!rc && e != hipSuccess)
            rc = mlp_set_error(BHSPARSE_HIP_ERROR, std::string("mlp f64: ") + hipGetErrorString(e));
    }
    (void)hipStreamDestroy(st);
    return rc;
}
catch (const std::bad_alloc &) {   // no exception crosses the C boundary
    return mlp_set_error(BHSPARSE_OUT_OF_MEMORY, "host allocation failed");
} catch (...) {
    return mlp_set_error(BHSPARSE_HIP_ERROR, "internal error");
}

} // extern "C"
