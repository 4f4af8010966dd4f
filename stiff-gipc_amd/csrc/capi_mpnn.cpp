// C ABI of the learned coarse level: network parameters, training sets over a
// level-1 graph, forward pass, loss gradient and Adam training.  No
// counterpart in the reference.
#include "capi_internal.hpp"
#include "mpnn.hpp"

#include <algorithm>
#include <numeric>
#include <random>
#include <vector>

struct sgipc_mpnn {
    sgipc_ctx* ctx;
    std::unique_ptr<sgipc::Mpnn> net;
};

struct sgipc_mpnn_data {
    sgipc_ctx* ctx;
    sgipc::CoarseGraph graph;
    sgipc::DevBuf<float> X;  // [S][3n][3]
    sgipc::DevBuf<float> q;  // [S][3n]
    sgipc::DevBuf<float> bX, bq;  // minibatch staging
    int count = 0;
    int cap = 0;
};

using namespace sgipc;

namespace {

void grow(sgipc_mpnn_data* d, int need)
{
    if (need <= d->cap) return;
    int cap = std::max(need, d->cap * 2 + 16);
    const long rows = 3L * d->graph.n;
    DevBuf<float> X2, q2;
    X2.reserve((size_t)cap * rows * 3);
    q2.reserve((size_t)cap * rows);
    hipStream_t s = d->ctx->c.stream;
    if (d->count > 0) {
        SGIPC_HIP(hipMemcpyAsync(X2.get(), d->X.get(), sizeof(float) * d->count * rows * 3, hipMemcpyDeviceToDevice, s));
        SGIPC_HIP(hipMemcpyAsync(q2.get(), d->q.get(), sizeof(float) * d->count * rows, hipMemcpyDeviceToDevice, s));
        SGIPC_HIP(hipStreamSynchronize(s));
    }
    std::swap(d->X.p, X2.p);
    std::swap(d->X.n, X2.n);
    std::swap(d->q.p, q2.p);
    std::swap(d->q.n, q2.n);
    d->cap = cap;
}

}  // namespace

namespace sgipc {

// Append samples that live on the device (used by the system's collector).
void mpnn_data_append(sgipc_mpnn_data* d, const float* X, const float* q, int n_samples)
{
    grow(d, d->count + n_samples);
    const long rows = 3L * d->graph.n;
    hipStream_t s = d->ctx->c.stream;
    SGIPC_HIP(hipMemcpyAsync(d->X.get() + (size_t)d->count * rows * 3, X, sizeof(float) * n_samples * rows * 3,
                             hipMemcpyDeviceToDevice, s));
    SGIPC_HIP(hipMemcpyAsync(d->q.get() + (size_t)d->count * rows, q, sizeof(float) * n_samples * rows,
                             hipMemcpyDeviceToDevice, s));
    d->count += n_samples;
}

sgipc_mpnn_data* mpnn_data_new(sgipc_ctx* ctx)
{
    auto* d = new sgipc_mpnn_data();
    d->ctx = ctx;
    return d;
}

CoarseGraph& mpnn_data_graph(sgipc_mpnn_data* d) { return d->graph; }

Mpnn* mpnn_of(sgipc_mpnn* h) { return h ? h->net.get() : nullptr; }

}  // namespace sgipc

extern "C" {

int sgipc_mpnn_create(sgipc_ctx* ctx, int layers, int linear, sgipc_mpnn** out)
{
    if (!ctx || !out) return SGIPC_BAD_ARG;
    *out = nullptr;
    if (layers < 1 || layers > kMpnnMaxLayers) return SGIPC_BAD_ARG;
    auto* h = new (std::nothrow) sgipc_mpnn();
    if (!h) return SGIPC_HIP_ERROR;
    h->ctx = ctx;
    int st = guarded(ctx, [&] { h->net = std::make_unique<Mpnn>(&ctx->c, layers, linear); });
    if (st != SGIPC_OK) {
        delete h;
        return st;
    }
    *out = h;
    return SGIPC_OK;
}

int sgipc_mpnn_destroy(sgipc_mpnn* h)
{
    if (!h) return SGIPC_BAD_ARG;
    int st = guarded(h->ctx, [&] { h->net.reset(); });
    delete h;
    return st;
}

int sgipc_mpnn_num_params(const sgipc_mpnn* h) { return h ? h->net->n_params() : -1; }

int sgipc_mpnn_set_params(sgipc_mpnn* h, const float* host, int n)
{
    if (!h || !host) return SGIPC_BAD_ARG;
    if (n != h->net->n_params()) return SGIPC_BAD_ARG;
    return guarded(h->ctx, [&] { h->net->set_params(host); });
}

int sgipc_mpnn_get_params(const sgipc_mpnn* h, float* host, int n)
{
    if (!h || !host) return SGIPC_BAD_ARG;
    if (n != h->net->n_params()) return SGIPC_BAD_ARG;
    return guarded(h->ctx, [&] { h->net->get_params(host); });
}

int sgipc_mpnn_init_params(sgipc_mpnn* h, unsigned long long seed)
{
    if (!h) return SGIPC_BAD_ARG;
    return guarded(h->ctx, [&] { h->net->init_random(seed); });
}

int sgipc_mpnn_data_create(sgipc_ctx* ctx, int n_nodes, const double* vals, const int* rows, const int* cols, int nb,
                           sgipc_mpnn_data** out)
{
    if (!ctx || !out) return SGIPC_BAD_ARG;
    *out = nullptr;
    if (n_nodes < 1 || nb < 1) return SGIPC_EMPTY;
    if (!vals || !rows || !cols) return SGIPC_BAD_ARG;
    auto* d = mpnn_data_new(ctx);
    int st = guarded(ctx, [&] {
        d->graph.build(ctx->c.stream, vals, rows, cols, nb, n_nodes);
        SGIPC_HIP(hipStreamSynchronize(ctx->c.stream));
    });
    if (st != SGIPC_OK) {
        delete d;
        return st;
    }
    *out = d;
    return SGIPC_OK;
}

int sgipc_mpnn_data_add(sgipc_mpnn_data* d, const float* X, const float* q, int n_samples)
{
    if (!d || !X || !q || n_samples < 1) return SGIPC_BAD_ARG;
    return guarded(d->ctx, [&] {
        mpnn_data_append(d, X, q, n_samples);
        SGIPC_HIP(hipStreamSynchronize(d->ctx->c.stream));
    });
}

int sgipc_mpnn_data_size(const sgipc_mpnn_data* d) { return d ? d->count : -1; }

int sgipc_mpnn_data_nodes(const sgipc_mpnn_data* d) { return d ? d->graph.n : -1; }

int sgipc_mpnn_data_samples(const sgipc_mpnn_data* d, int first, int count, float* X_host, float* q_host)
{
    if (!d || first < 0 || count < 0 || first + count > d->count) return SGIPC_BAD_ARG;
    return guarded(d->ctx, [&] {
        const long rows = 3L * d->graph.n;
        hipStream_t s = d->ctx->c.stream;
        if (X_host)
            SGIPC_HIP(hipMemcpyAsync(X_host, d->X.get() + (size_t)first * rows * 3, sizeof(float) * count * rows * 3,
                                     hipMemcpyDeviceToHost, s));
        if (q_host)
            SGIPC_HIP(hipMemcpyAsync(q_host, d->q.get() + (size_t)first * rows, sizeof(float) * count * rows,
                                     hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
    });
}

int sgipc_mpnn_data_graph(const sgipc_mpnn_data* d, int* row_ptr, int* col, float* blk, float* scale)
{
    if (!d) return SGIPC_BAD_ARG;
    return guarded(d->ctx, [&] {
        const CoarseGraph& g = d->graph;
        hipStream_t s = d->ctx->c.stream;
        if (row_ptr)
            SGIPC_HIP(hipMemcpyAsync(row_ptr, g.row_ptr.get(), sizeof(int) * (g.n + 1), hipMemcpyDeviceToHost, s));
        if (col) SGIPC_HIP(hipMemcpyAsync(col, g.col.get(), sizeof(int) * g.nnz, hipMemcpyDeviceToHost, s));
        if (blk) SGIPC_HIP(hipMemcpyAsync(blk, g.blk.get(), sizeof(float) * 9L * g.nnz, hipMemcpyDeviceToHost, s));
        if (scale) SGIPC_HIP(hipMemcpyAsync(scale, g.s.get(), sizeof(float) * g.n, hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
    });
}

int sgipc_mpnn_data_graph_nnz(const sgipc_mpnn_data* d) { return d ? d->graph.nnz : -1; }

int sgipc_mpnn_data_destroy(sgipc_mpnn_data* d)
{
    if (!d) return SGIPC_BAD_ARG;
    sgipc_ctx* ctx = d->ctx;
    int st = guarded(ctx, [&] { delete d; });
    return st;
}

int sgipc_mpnn_forward(sgipc_mpnn* h, const sgipc_mpnn_data* d, int first, int count, float* chat)
{
    if (!h || !d || !chat || first < 0 || count < 1 || first + count > d->count) return SGIPC_BAD_ARG;
    return guarded(h->ctx, [&] {
        const long rows = 3L * d->graph.n;
        h->net->forward(h->ctx->c.stream, d->graph, d->X.get() + (size_t)first * rows * 3, count, chat, nullptr);
        SGIPC_HIP(hipStreamSynchronize(h->ctx->c.stream));
    });
}

int sgipc_mpnn_loss_grad(sgipc_mpnn* h, const sgipc_mpnn_data* d, int first, int count, double* loss,
                         float* grad_host)
{
    if (!h || !d || first < 0 || count < 1 || first + count > d->count) return SGIPC_BAD_ARG;
    return guarded(h->ctx, [&] {
        const long rows = 3L * d->graph.n;
        hipStream_t s = h->ctx->c.stream;
        Mpnn& net = *h->net;
        DevBuf<float> g;
        g.reserve(net.n_params());
        const double l = net.loss_grad(s, d->graph, d->X.get() + (size_t)first * rows * 3,
                                       d->q.get() + (size_t)first * rows, count, g.get());
        if (loss) *loss = l;
        if (grad_host)
            SGIPC_HIP(hipMemcpyAsync(grad_host, g.get(), sizeof(float) * net.n_params(), hipMemcpyDeviceToHost, s));
        SGIPC_HIP(hipStreamSynchronize(s));
    });
}

int sgipc_mpnn_train(sgipc_mpnn* h, sgipc_mpnn_data* d, int epochs, int batch, float lr, unsigned seed,
                     float* epoch_loss)
{
    if (!h || !d || epochs < 1 || batch < 1 || !(lr > 0.f)) return SGIPC_BAD_ARG;
    if (d->count < 1) return SGIPC_EMPTY;
    return guarded(h->ctx, [&] {
        Mpnn& net = *h->net;
        hipStream_t s = h->ctx->c.stream;
        const long rows = 3L * d->graph.n;
        const int bs = std::min(batch, d->count);
        d->bX.reserve((size_t)bs * rows * 3);
        d->bq.reserve((size_t)bs * rows);
        std::mt19937 rng(seed);
        std::vector<int> order(d->count);
        for (int e = 0; e < epochs; ++e) {
            std::iota(order.begin(), order.end(), 0);
            std::shuffle(order.begin(), order.end(), rng);
            double acc = 0.0;
            int nbatch = 0;
            for (int b0 = 0; b0 < d->count; b0 += bs) {
                const int cnt = std::min(bs, d->count - b0);
                for (int k = 0; k < cnt; ++k) {
                    const int id = order[b0 + k];
                    SGIPC_HIP(hipMemcpyAsync(d->bX.get() + (size_t)k * rows * 3, d->X.get() + (size_t)id * rows * 3,
                                             sizeof(float) * rows * 3, hipMemcpyDeviceToDevice, s));
                    SGIPC_HIP(hipMemcpyAsync(d->bq.get() + (size_t)k * rows, d->q.get() + (size_t)id * rows,
                                             sizeof(float) * rows, hipMemcpyDeviceToDevice, s));
                }
                acc += net.loss_grad(s, d->graph, d->bX.get(), d->bq.get(), cnt, net.grad_buf());
                net.adam_step(s, lr, 0.9f, 0.999f, 1e-8f);
                ++nbatch;
            }
            if (epoch_loss) epoch_loss[e] = (float)(acc / nbatch);
        }
        SGIPC_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
