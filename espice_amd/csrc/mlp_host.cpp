// Host side of the learned warm start (include/espice_lanes.h, "Learned warm
// start"): parameter layout, initialization, the training loop and the
// device buffers. Every product, loss, gradient and update runs in the
// kernels of mlp.hip; the host only shuffles row indices and sequences
// launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/espice_lanes.h"
#include "esl_host.h"
#include "esl_mlp_internal.h"
#include "esl_warm.h"

using std::vector;
using u32 = uint32_t;
using eslh::DevBuf;

namespace {

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
double unit(uint64_t &s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct EslMlp {
    vector<u32> dims;                 // n_in, hidden..., n_out
    vector<size_t> w_off, b_off;
    size_t n_params = 0;
    DevBuf p, pb, grad, m, v, is_weight;
    DevBuf in_mean, in_scale, out_mean, out_scale;
    // work
    u32 cap_rows = 0;
    DevBuf X0, dZf, dZb, dHf, row_loss, loss_acc;
    vector<DevBuf *> Z, H;            // per layer: FP32 pre-activation, BF16 activation
    DevBuf inputs, targets, T, idx;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t adam_t = 0;

    u32 layers() const { return (u32)dims.size() - 1; }
    u32 n_in() const { return dims.front(); }
    u32 n_out() const { return dims.back(); }
    u32 widest() const { return *std::max_element(dims.begin(), dims.end()); }
    ~EslMlp()
    {
        for (auto *b : Z) delete b;
        for (auto *b : H) delete b;
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

esl_status work(EslMlp &n, u32 rows)
{
    if (rows <= n.cap_rows) return ESL_OK;
    const size_t w = n.widest();
    HIPCHK(n.X0.ensure((size_t)rows * n.n_in() * 2));
    HIPCHK(n.dZf.ensure((size_t)rows * w * 4));
    HIPCHK(n.dZb.ensure((size_t)rows * w * 2));
    HIPCHK(n.dHf.ensure((size_t)rows * w * 4));
    HIPCHK(n.row_loss.ensure((size_t)rows * 4));
    for (u32 l = 0; l < n.layers(); l++) {
        HIPCHK(n.Z[l]->ensure((size_t)rows * n.dims[l + 1] * 4));
        HIPCHK(n.H[l]->ensure((size_t)rows * n.dims[l + 1] * 2));
    }
    n.cap_rows = rows;
    return ESL_OK;
}

esl::GemmArgs gemm(const uint16_t *A, int64_t sa_m, int64_t sa_k, const uint16_t *B, int64_t sb_k, int64_t sb_n,
                   u32 M, u32 N, u32 K)
{
    esl::GemmArgs g{};
    g.A = A; g.sa_m = sa_m; g.sa_k = sa_k;
    g.B = B; g.sb_k = sb_k; g.sb_n = sb_n;
    g.M = M; g.N = N; g.K = K;
    return g;
}

const uint16_t *layer_input(const EslMlp &n, u32 l)
{
    return l == 0 ? n.X0.as<uint16_t>() : n.H[l - 1]->as<uint16_t>();
}

// Z_l = X_l W_l^T + b_l; H_l = silu(Z_l) (hidden layers).
esl_status forward(EslMlp &n, u32 rows, hipStream_t s)
{
    for (u32 l = 0; l < n.layers(); l++) {
        const u32 in = n.dims[l], out = n.dims[l + 1];
        const bool hidden = l + 1 < n.layers();
        esl::GemmArgs g = gemm(layer_input(n, l), in, 1, n.pb.as<uint16_t>() + n.w_off[l], 1, in, rows, out, in);
        g.bias = n.p.as<float>() + n.b_off[l];
        g.Cf = n.Z[l]->as<float>();
        g.sc_m = out; g.sc_n = 1;
        if (hidden) {
            g.Cb = n.H[l]->as<uint16_t>();
            g.sd_m = out; g.sd_n = 1;
            g.act = 1;
        }
        HIPCHK(esl_mlp_launch_gemm(&g, s));
    }
    return ESL_OK;
}

// Gradients of the loss whose output gradient is in dZf/dZb.
esl_status backward(EslMlp &n, u32 rows, hipStream_t s)
{
    for (u32 l = n.layers(); l > 0;) {
        l--;
        const u32 in = n.dims[l], out = n.dims[l + 1];
        // dW_l[j][k] = sum_r dZ[r][j] X_l[r][k]
        esl::GemmArgs g = gemm(n.dZb.as<uint16_t>(), 1, out, layer_input(n, l), in, 1, out, in, rows);
        g.Cf = n.grad.as<float>() + n.w_off[l];
        g.sc_m = in; g.sc_n = 1;
        HIPCHK(esl_mlp_launch_gemm(&g, s));
        HIPCHK(esl_mlp_launch_bias_grad(n.dZf.as<float>(), rows, out, n.grad.as<float>() + n.b_off[l], s));
        if (l == 0) break;
        // dH_{l-1}[r][k] = sum_j dZ[r][j] W_l[j][k]; dZ_{l-1} = dH silu'(Z_{l-1})
        esl::GemmArgs h = gemm(n.dZb.as<uint16_t>(), out, 1, n.pb.as<uint16_t>() + n.w_off[l], in, 1, rows, in, out);
        h.Cf = n.dHf.as<float>();
        h.sc_m = in; h.sc_n = 1;
        HIPCHK(esl_mlp_launch_gemm(&h, s));
        HIPCHK(esl_mlp_launch_act_back(n.dHf.as<float>(), n.Z[l - 1]->as<float>(), rows * in, n.dZf.as<float>(),
                                       n.dZb.as<uint16_t>(), s));
    }
    return ESL_OK;
}

esl_status upload_rows(EslMlp &n, const double *inputs, const double *targets, size_t rows)
{
    hipStream_t s = n.stream;
    HIPCHK(n.inputs.ensure(rows * n.n_in() * 8));
    HIPCHK(hipMemcpyAsync(n.inputs.p, inputs, rows * n.n_in() * 8, hipMemcpyHostToDevice, s));
    if (targets) {
        const size_t w = (size_t)n.n_out() + 1;
        HIPCHK(n.targets.ensure(rows * w * 8));
        HIPCHK(n.T.ensure(rows * n.n_out() * 4));
        HIPCHK(hipMemcpyAsync(n.targets.p, targets, rows * w * 8, hipMemcpyHostToDevice, s));
        HIPCHK(esl_mlp_launch_targets(n.targets.as<double>(), (u32)rows, n.n_out(), n.out_mean.as<double>(),
                                      n.out_scale.as<double>(), n.T.as<float>(), s));
    }
    return ESL_OK;
}

bool finite_rows(const double *a, size_t count)
{
    for (size_t i = 0; i < count; i++)
        if (!std::isfinite(a[i])) return false;
    return true;
}

}  // namespace

namespace eslw {

esl_status predict_device(EslMlp &n, const double *d_inputs, u32 rows, double *d_x, hipStream_t s)
{
    esl_status st = work(n, rows);
    if (st != ESL_OK) return st;
    HIPCHK(esl_mlp_launch_inputs(d_inputs, nullptr, rows, n.n_in(), n.in_mean.as<double>(), n.in_scale.as<double>(),
                                 n.X0.as<uint16_t>(), s));
    st = forward(n, rows, s);
    if (st != ESL_OK) return st;
    HIPCHK(esl_mlp_launch_outputs(n.Z[n.layers() - 1]->as<float>(), rows, n.n_out(), n.out_mean.as<double>(),
                                  n.out_scale.as<double>(), d_x, s));
    return ESL_OK;
}

EslMlp *lookup_mlp(esl_mlp *h) { return static_cast<EslMlp *>(eslh::lookup(h, 3)); }
u32 mlp_n_in(const EslMlp &n) { return n.n_in(); }
u32 mlp_n_out(const EslMlp &n) { return n.n_out(); }

}  // namespace eslw

extern "C" {

void esl_mlp_default_train_options(esl_mlp_train_options *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->struct_size = sizeof(*o);
    o->epochs = 40;
    o->batch = 512;
    o->lr = 2e-3f;
    o->lr_final = 2e-5f;
    o->beta1 = 0.9f;
    o->beta2 = 0.999f;
    o->eps = 1e-8f;
    o->weight_decay = 0.0f;
    o->seed = 1;
}

esl_status esl_mlp_create(uint32_t n_in, uint32_t n_hidden, uint32_t n_hidden_layers, uint32_t n_out, uint64_t seed,
                          esl_mlp **out)
{
    if (!out) return ESL_INVALID_ARGUMENT;
    *out = nullptr;
    if (n_in == 0 || n_hidden == 0 || n_out == 0 || n_hidden_layers == 0 || n_hidden_layers > 16) return ESL_INVALID_ARGUMENT;
    if (n_in > 65536 || n_hidden > 65536 || n_out > 65536) return ESL_INVALID_ARGUMENT;
    if (!eslh::device_present()) return ESL_DEVICE_UNAVAILABLE;
    auto *n = new EslMlp();
    n->dims.push_back(n_in);
    for (u32 i = 0; i < n_hidden_layers; i++) n->dims.push_back(n_hidden);
    n->dims.push_back(n_out);
    size_t off = 0;
    for (u32 l = 0; l < n->layers(); l++) {
        n->w_off.push_back(off);
        off += (size_t)n->dims[l] * n->dims[l + 1];
        n->b_off.push_back(off);
        off += n->dims[l + 1];
    }
    n->n_params = off;
    vector<float> p(off, 0.f);
    vector<uint8_t> isw(off, 0);
    uint64_t s = seed;
    for (u32 l = 0; l < n->layers(); l++) {
        const double a = std::sqrt(6.0 / (n->dims[l] + n->dims[l + 1]));
        for (size_t i = 0; i < (size_t)n->dims[l] * n->dims[l + 1]; i++) {
            p[n->w_off[l] + i] = (float)((2 * unit(s) - 1) * a);
            isw[n->w_off[l] + i] = 1;
        }
    }
    auto fail = [&](esl_status st) {
        delete n;
        return st;
    };
    auto hip = [&](hipError_t e, const char *w) { return e == hipSuccess ? ESL_OK : eslh::fail_hip(e, w); };
    if (hip(hipStreamCreate(&n->stream), "stream") != ESL_OK) return fail(ESL_DEVICE_UNAVAILABLE);
    if (hip(hipEventCreate(&n->ev0), "event") != ESL_OK || hip(hipEventCreate(&n->ev1), "event") != ESL_OK)
        return fail(ESL_DEVICE_UNAVAILABLE);
    for (u32 l = 0; l < n->layers(); l++) {
        n->Z.push_back(new DevBuf());
        n->H.push_back(new DevBuf());
    }
    if (n->p.ensure(off * 4) != hipSuccess || n->pb.ensure(off * 2) != hipSuccess ||
        n->grad.ensure(off * 4) != hipSuccess || n->m.ensure(off * 4) != hipSuccess ||
        n->v.ensure(off * 4) != hipSuccess || n->is_weight.ensure(off) != hipSuccess ||
        n->in_mean.ensure(n_in * 8) != hipSuccess || n->in_scale.ensure(n_in * 8) != hipSuccess ||
        n->out_mean.ensure(n_out * 8) != hipSuccess || n->out_scale.ensure(n_out * 8) != hipSuccess ||
        n->loss_acc.ensure(16) != hipSuccess)
        return fail(ESL_OUT_OF_MEMORY);
    vector<double> zi(n_in, 0.0), oi(n_in, 1.0), zo(n_out, 0.0), oo(n_out, 1.0);
    if (hipMemcpy(n->p.p, p.data(), off * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->is_weight.p, isw.data(), off, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m.p, 0, off * 4) != hipSuccess || hipMemset(n->v.p, 0, off * 4) != hipSuccess ||
        hipMemcpy(n->in_mean.p, zi.data(), n_in * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->in_scale.p, oi.data(), n_in * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->out_mean.p, zo.data(), n_out * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(n->out_scale.p, oo.data(), n_out * 8, hipMemcpyHostToDevice) != hipSuccess)
        return fail(ESL_DEVICE_UNAVAILABLE);
    if (esl_mlp_launch_to_bf16(n->p.as<float>(), (u32)off, n->pb.as<uint16_t>(), n->stream) != hipSuccess ||
        hipStreamSynchronize(n->stream) != hipSuccess)
        return fail(ESL_DEVICE_UNAVAILABLE);
    *out = static_cast<esl_mlp *>(eslh::make_handle(n, 3));
    return ESL_OK;
}

esl_status esl_mlp_destroy(esl_mlp *h)
{
    if (!h) return ESL_OK;
    auto *n = static_cast<EslMlp *>(eslh::release(h, 3));
    if (!n) return ESL_INVALID_HANDLE;
    delete n;
    return ESL_OK;
}

esl_status esl_mlp_param_count(esl_mlp *h, size_t *count)
{
    auto *n = eslw::lookup_mlp(h);
    if (!n) return ESL_INVALID_HANDLE;
    if (!count) return ESL_INVALID_ARGUMENT;
    *count = n->n_params;
    return ESL_OK;
}

esl_status esl_mlp_get_params(esl_mlp *h, float *params, size_t count)
{
    auto *n = eslw::lookup_mlp(h);
    if (!n) return ESL_INVALID_HANDLE;
    if (!params) return ESL_INVALID_ARGUMENT;
    if (count < n->n_params) return ESL_BUFFER_TOO_SMALL;
    HIPCHK(hipStreamSynchronize(n->stream));
    HIPCHK(hipMemcpy(params, n->p.p, n->n_params * 4, hipMemcpyDeviceToHost));
    return ESL_OK;
}

esl_status esl_mlp_set_params(esl_mlp *h, const float *params, size_t count)
{
    auto *n = eslw::lookup_mlp(h);
    if (!n) return ESL_INVALID_HANDLE;
    if (!params || count != n->n_params) return ESL_INVALID_ARGUMENT;
    for (size_t i = 0; i < count; i++)
        if (!std::isfinite(params[i])) return ESL_INVALID_ARGUMENT;
    HIPCHK(hipStreamSynchronize(n->stream));
    HIPCHK(hipMemcpy(n->p.p, params, count * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(n->m.p, 0, count * 4));
    HIPCHK(hipMemset(n->v.p, 0, count * 4));
    n->adam_t = 0;
    HIPCHK(esl_mlp_launch_to_bf16(n->p.as<float>(), (u32)count, n->pb.as<uint16_t>(), n->stream));
    HIPCHK(hipStreamSynchronize(n->stream));
    return ESL_OK;
}

esl_status esl_mlp_set_scaling(esl_mlp *h, const double *in_mean, const double *in_scale, const double *out_mean,
                               const double *out_scale)
{
    auto *n = eslw::lookup_mlp(h);
    if (!n) return ESL_INVALID_HANDLE;
    if (!in_mean || !in_scale || !out_mean || !out_scale) return ESL_INVALID_ARGUMENT;
    for (u32 i = 0; i < n->n_in(); i++)
        if (!std::isfinite(in_mean[i]) || !(in_scale[i] > 0) || !std::isfinite(in_scale[i])) return ESL_INVALID_ARGUMENT;
    for (u32 i = 0; i < n->n_out(); i++)
        if (!std::isfinite(out_mean[i]) || !(out_scale[i] > 0) || !std::isfinite(out_scale[i])) return ESL_INVALID_ARGUMENT;
    HIPCHK(hipStreamSynchronize(n->stream));
    HIPCHK(hipMemcpy(n->in_mean.p, in_mean, n->n_in() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(n->in_scale.p, in_scale, n->n_in() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(n->out_mean.p, out_mean, n->n_out() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(n->out_scale.p, out_scale, n->n_out() * 8, hipMemcpyHostToDevice));
    return ESL_OK;
}

esl_status esl_mlp_gradient(esl_mlp *h, const double *inputs, const double *targets, size_t rows, float *loss,
                            float *grads)
{
    auto *n = eslw::lookup_mlp(h);
    if (!n) return ESL_INVALID_HANDLE;
    if (!inputs || !targets || rows == 0 || rows > (1u << 24)) return ESL_INVALID_ARGUMENT;
    if (!finite_rows(inputs, rows * n->n_in()) || !finite_rows(targets, rows * (n->n_out() + 1)))
        return ESL_INVALID_ARGUMENT;
    hipStream_t s = n->stream;
    esl_status st = work(*n, (u32)rows);
    if (st == ESL_OK) st = upload_rows(*n, inputs, targets, rows);
    if (st != ESL_OK) return st;
    HIPCHK(hipMemsetAsync(n->loss_acc.p, 0, 16, s));
    HIPCHK(esl_mlp_launch_inputs(n->inputs.as<double>(), nullptr, (u32)rows, n->n_in(), n->in_mean.as<double>(),
                                 n->in_scale.as<double>(), n->X0.as<uint16_t>(), s));
    st = forward(*n, (u32)rows, s);
    if (st != ESL_OK) return st;
    HIPCHK(esl_mlp_launch_loss(n->Z[n->layers() - 1]->as<float>(), n->T.as<float>(), nullptr, (u32)rows, n->n_out(),
                               n->dZf.as<float>(), n->dZb.as<uint16_t>(), n->row_loss.as<float>(),
                               n->loss_acc.as<double>(), s));
    st = backward(*n, (u32)rows, s);
    if (st != ESL_OK) return st;
    double acc[2];
    HIPCHK(hipMemcpyAsync(acc, n->loss_acc.p, 16, hipMemcpyDeviceToHost, s));
    if (grads) HIPCHK(hipMemcpyAsync(grads, n->grad.p, n->n_params * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (loss) *loss = (float)acc[0];
    return ESL_OK;
}

esl_status esl_mlp_train(esl_mlp *h, const double *inputs, const double *targets, size_t rows,
                         const esl_mlp_train_options *opt, float *loss_per_epoch)
{
    auto *n = eslw::lookup_mlp(h);
    if (!n) return ESL_INVALID_HANDLE;
    esl_mlp_train_options o;
    esl_mlp_default_train_options(&o);
    if (opt) {
        if (opt->struct_size < sizeof(o)) return ESL_ABI_MISMATCH;
        o = *opt;
    }
    if (!inputs || !targets || rows == 0 || rows > (1u << 24)) return ESL_INVALID_ARGUMENT;
    if (o.batch == 0 || o.epochs == 0 || !(o.lr > 0)) return ESL_INVALID_ARGUMENT;
    if (!finite_rows(inputs, rows * n->n_in()) || !finite_rows(targets, rows * (n->n_out() + 1)))
        return ESL_INVALID_ARGUMENT;
    const u32 batch = (u32)std::min<size_t>(o.batch, rows);
    const size_t steps = rows / batch;
    hipStream_t s = n->stream;
    esl_status st = work(*n, batch);
    if (st == ESL_OK) st = upload_rows(*n, inputs, targets, rows);
    if (st != ESL_OK) return st;
    HIPCHK(n->idx.ensure(rows * 4));
    vector<u32> perm(rows);
    uint64_t rs = o.seed;
    const size_t total = (size_t)o.epochs * steps;
    size_t k = 0;
    for (u32 ep = 0; ep < o.epochs; ep++) {
        for (size_t i = 0; i < rows; i++) perm[i] = (u32)i;
        for (size_t i = rows; i > 1; i--) std::swap(perm[i - 1], perm[splitmix(rs) % i]);
        HIPCHK(hipMemcpyAsync(n->idx.p, perm.data(), rows * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemsetAsync(n->loss_acc.p, 0, 16, s));
        for (size_t b = 0; b < steps; b++, k++) {
            const u32 *ix = n->idx.as<u32>() + b * batch;
            HIPCHK(esl_mlp_launch_inputs(n->inputs.as<double>(), ix, batch, n->n_in(), n->in_mean.as<double>(),
                                         n->in_scale.as<double>(), n->X0.as<uint16_t>(), s));
            st = forward(*n, batch, s);
            if (st != ESL_OK) return st;
            HIPCHK(esl_mlp_launch_loss(n->Z[n->layers() - 1]->as<float>(), n->T.as<float>(), ix, batch, n->n_out(),
                                       n->dZf.as<float>(), n->dZb.as<uint16_t>(), n->row_loss.as<float>(),
                                       n->loss_acc.as<double>(), s));
            st = backward(*n, batch, s);
            if (st != ESL_OK) return st;
            n->adam_t++;
            const double frac = total > 1 ? (double)k / (double)(total - 1) : 1.0;
            esl::AdamStep a{};
            a.lr = (float)(o.lr_final + 0.5 * (o.lr - o.lr_final) * (1.0 + std::cos(M_PI * frac)));
            a.beta1 = o.beta1;
            a.beta2 = o.beta2;
            a.eps = o.eps;
            a.weight_decay = o.weight_decay;
            a.bc1 = (float)(1.0 - std::pow((double)o.beta1, (double)n->adam_t));
            a.bc2 = (float)(1.0 - std::pow((double)o.beta2, (double)n->adam_t));
            HIPCHK(esl_mlp_launch_adam(n->p.as<float>(), n->grad.as<float>(), n->m.as<float>(), n->v.as<float>(),
                                       (u32)n->n_params, &a, n->is_weight.as<uint8_t>(), n->pb.as<uint16_t>(), s));
        }
        double acc[2];
        HIPCHK(hipMemcpyAsync(acc, n->loss_acc.p, 16, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (loss_per_epoch) loss_per_epoch[ep] = (float)(acc[1] > 0 ? acc[0] / acc[1] : 0.0);
    }
    return ESL_OK;
}

esl_status esl_mlp_predict(esl_mlp *h, const double *inputs, size_t rows, double *out)
{
    auto *n = eslw::lookup_mlp(h);
    if (!n) return ESL_INVALID_HANDLE;
    if (!inputs || !out || rows == 0 || rows > (1u << 24)) return ESL_INVALID_ARGUMENT;
    if (!finite_rows(inputs, rows * n->n_in())) return ESL_INVALID_ARGUMENT;
    hipStream_t s = n->stream;
    esl_status st = upload_rows(*n, inputs, nullptr, rows);
    if (st != ESL_OK) return st;
    DevBuf x;
    HIPCHK(x.ensure(rows * (n->n_out() + 1) * 8));
    st = eslw::predict_device(*n, n->inputs.as<double>(), (u32)rows, x.as<double>(), s);
    if (st != ESL_OK) return st;
    HIPCHK(hipMemcpyAsync(out, x.p, rows * (n->n_out() + 1) * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return ESL_OK;
}

}  // extern "C"
