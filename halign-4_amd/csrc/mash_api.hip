// mash_api.hip -- host side of the sketches (kernels in mash.hip): the C ABI
// halign_mash_sketch / halign_mash_best and the device-resident forms the
// multi-reference keep-length run uses.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "host_common.h"
#include "mash_host.h"

namespace hal {

hipError_t mash_launch(int phase, int n, const char *seqs, const int64_t *off, const int64_t *key_off, int k,
                       int noncanonical, uint64_t seed_mix, uint64_t *keys, int64_t *nkeys, int size, uint64_t *out,
                       int32_t *count, hipStream_t stream);
hipError_t mash_select_launch(int nq, const uint64_t *qsk, const int32_t *qcnt, int qstride, const uint64_t *rsk,
                              const int32_t *rcnt, int rstride, int nr, int32_t *best, int32_t *inter, int32_t *den,
                              hipStream_t stream);

namespace {
constexpr int64_t kMaxKeys = 32 << 20;   // k-mer keys sorted per pass (256 MiB per key buffer)
}

int mash_sketch_to_device(halign_ctx *ctx, int n, const char *seqs, const int64_t *offsets, int k, int size,
                          bool noncanonical, uint32_t seed, DevBuf &sk, DevBuf &cnt)
{
    hipStream_t st = ctx->stream;
    HCHECK(sk.ensure((size_t)std::max(n, 1) * (size_t)std::max(size, 1) * 8), "alloc sketches");
    HCHECK(cnt.ensure((size_t)std::max(n, 1) * 4), "alloc sketch counts");
    if (n == 0) return HALIGN_OK;
    if (k <= 0 || k > 32 || size <= 0) {   // the reference returns empty sketches
        HCHECK(hipMemsetAsync(cnt.p, 0, (size_t)n * 4, st), "clear sketch counts");
        return HALIGN_OK;
    }
    uint64_t mix = seed;
    mix ^= (uint64_t)__builtin_bswap32((uint32_t)mix) << 32;
    for (int a = 0; a < n;) {
        // sequences [a, b): their k-mer key ranges fit one pass
        int b = a;
        int64_t keys = 0;
        std::vector<int64_t> koff{0}, off;
        while (b < n) {
            const int64_t len = offsets[b + 1] - offsets[b];
            const int64_t cap = std::max<int64_t>(len - k + 1, 0);
            if (b > a && keys + cap > kMaxKeys) break;
            keys += cap;
            koff.push_back(keys);
            ++b;
        }
        const int m = b - a;
        off.assign(offsets + a, offsets + b + 1);
        const int64_t base = off[0];
        for (auto &o : off) o -= base;
        const int64_t bytes = off[m];
        HCHECK(ctx->m_seqs.ensure((size_t)std::max<int64_t>(bytes, 1)), "alloc sketch input");
        HCHECK(ctx->m_off.ensure((size_t)(m + 1) * 8), "alloc");
        HCHECK(ctx->m_koff.ensure((size_t)(m + 1) * 8), "alloc");
        HCHECK(ctx->m_keys.ensure((size_t)std::max<int64_t>(keys, 1) * 8), "alloc keys");
        HCHECK(ctx->m_keys2.ensure((size_t)std::max<int64_t>(keys, 1) * 8), "alloc keys");
        HCHECK(ctx->m_nkeys.ensure((size_t)m * 8), "alloc");
        if (bytes) HCHECK(hipMemcpyAsync(ctx->m_seqs.p, seqs + base, (size_t)bytes, hipMemcpyHostToDevice, st), "upload");
        HCHECK(hipMemcpyAsync(ctx->m_off.p, off.data(), (size_t)(m + 1) * 8, hipMemcpyHostToDevice, st), "upload");
        HCHECK(hipMemcpyAsync(ctx->m_koff.p, koff.data(), (size_t)(m + 1) * 8, hipMemcpyHostToDevice, st), "upload");
        HCHECK(mash_launch(0, m, ctx->m_seqs.as<char>(), ctx->m_off.as<int64_t>(), ctx->m_koff.as<int64_t>(), k,
                           noncanonical ? 1 : 0, mix, ctx->m_keys.as<uint64_t>(), ctx->m_nkeys.as<int64_t>(), 0,
                           nullptr, nullptr, st),
               "launch mash_kmers");
        if (keys > 0) {
            // ranges [koff[i], koff[i] + nkeys[i]) hold the keys; a range is sorted
            // whole, so its tail is filled with the largest key first
            size_t tmp = 0;
            const int64_t *kb = ctx->m_koff.as<int64_t>();
            HCHECK(rocprim::segmented_radix_sort_keys(nullptr, tmp, ctx->m_keys.as<uint64_t>(),
                                                      ctx->m_keys2.as<uint64_t>(), (unsigned int)keys,
                                                      (unsigned int)m, kb, kb + 1, 0, 64, st),
                   "sort storage");
            HCHECK(ctx->m_tmp.ensure(std::max<size_t>(tmp, 1)), "alloc sort storage");
            HCHECK(rocprim::segmented_radix_sort_keys(ctx->m_tmp.p, tmp, ctx->m_keys.as<uint64_t>(),
                                                      ctx->m_keys2.as<uint64_t>(), (unsigned int)keys,
                                                      (unsigned int)m, kb, kb + 1, 0, 64, st),
                   "sort k-mer hashes");
        }
        HCHECK(mash_launch(1, m, nullptr, nullptr, ctx->m_koff.as<int64_t>(), k, 0, 0, ctx->m_keys2.as<uint64_t>(),
                           ctx->m_nkeys.as<int64_t>(), size, sk.as<uint64_t>() + (int64_t)a * size,
                           cnt.as<int32_t>() + a, st),
               "launch mash_take");
        HCHECK(hipStreamSynchronize(st), "sketch pass");   // the pass's buffers are reused by the next
        a = b;
    }
    return HALIGN_OK;
}

int mash_select_device(halign_ctx *ctx, int nq, const DevBuf &qsk, const DevBuf &qcnt, int qstride, int nr,
                       const DevBuf &rsk, const DevBuf &rcnt, int rstride, std::vector<int32_t> &best,
                       std::vector<int32_t> &num, std::vector<int32_t> &den)
{
    best.assign((size_t)nq, -1);
    num.assign((size_t)nq, 0);
    den.assign((size_t)nq, 1);
    if (nq == 0) return HALIGN_OK;
    hipStream_t st = ctx->stream;
    HCHECK(ctx->m_sel.ensure((size_t)nq * 12), "alloc selection");
    int32_t *d = ctx->m_sel.as<int32_t>();
    HCHECK(mash_select_launch(nq, qsk.as<uint64_t>(), qcnt.as<int32_t>(), qstride, rsk.as<uint64_t>(),
                              rcnt.as<int32_t>(), rstride, nr, d, d + nq, d + 2 * nq, st),
           "launch mash_select");
    HCHECK(hipMemcpyAsync(best.data(), d, (size_t)nq * 4, hipMemcpyDeviceToHost, st), "download");
    HCHECK(hipMemcpyAsync(num.data(), d + nq, (size_t)nq * 4, hipMemcpyDeviceToHost, st), "download");
    HCHECK(hipMemcpyAsync(den.data(), d + 2 * nq, (size_t)nq * 4, hipMemcpyDeviceToHost, st), "download");
    HCHECK(hipStreamSynchronize(st), "selection");
    return HALIGN_OK;
}

}  // namespace hal

using namespace hal;

extern "C" {

int halign_mash_sketch(halign_ctx *ctx, int32_t n, const char *seqs, const int64_t *offsets, int32_t k,
                       int32_t sketch_size, int32_t non_canonical, uint32_t seed, uint64_t *hashes_out, int64_t cap,
                       int64_t *sketch_offsets)
{
    if (!ctx || n < 0 || !sketch_offsets || (n > 0 && (!seqs || !offsets))) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    sketch_offsets[0] = 0;
    if (n == 0) return HALIGN_OK;
    const int size = std::max(sketch_size, 0);
    DevBuf sk, cnt;
    const int rc = mash_sketch_to_device(ctx, n, seqs, offsets, k, size, non_canonical != 0, seed, sk, cnt);
    if (rc != HALIGN_OK) {
        sk.release();
        cnt.release();
        return rc;
    }
    std::vector<int32_t> c((size_t)n);
    std::vector<uint64_t> h((size_t)n * (size_t)std::max(size, 1));
    hipError_t e = hipMemcpy(c.data(), cnt.p, (size_t)n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && size > 0) e = hipMemcpy(h.data(), sk.p, (size_t)n * size * 8, hipMemcpyDeviceToHost);
    sk.release();
    cnt.release();
    if (e != hipSuccess) return fail(ctx, HALIGN_E_HIP, "download sketches", e);
    int64_t o = 0;
    for (int i = 0; i < n; ++i) {
        if (o + c[i] > cap) return fail(ctx, HALIGN_E_CAPACITY, "sketch output too small");
        if (hashes_out) std::copy(h.begin() + (int64_t)i * size, h.begin() + (int64_t)i * size + c[i], hashes_out + o);
        o += c[i];
        sketch_offsets[i + 1] = o;
    }
    return HALIGN_OK;
}

int halign_mash_best(halign_ctx *ctx, int32_t nq, const uint64_t *q_hashes, const int64_t *q_offsets, int32_t nr,
                     const uint64_t *r_hashes, const int64_t *r_offsets, int32_t *best, double *jaccard)
{
    if (!ctx || nq < 0 || nr <= 0 || !best || !r_offsets || (nq > 0 && !q_offsets)) return HALIGN_E_ARG;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HCHECK(hipSetDevice(ctx->device), "set device");
    if (nq == 0) return HALIGN_OK;
    // fixed-stride device layout of both sketch sets
    auto layout = [](int n, const uint64_t *h, const int64_t *off, int &stride, std::vector<uint64_t> &flat,
                     std::vector<int32_t> &cnt) {
        stride = 1;
        for (int i = 0; i < n; ++i) stride = std::max<int>(stride, (int)(off[i + 1] - off[i]));
        flat.assign((size_t)n * stride, 0);
        cnt.resize((size_t)n);
        for (int i = 0; i < n; ++i) {
            cnt[i] = (int32_t)(off[i + 1] - off[i]);
            std::copy(h + off[i], h + off[i + 1], flat.begin() + (int64_t)i * stride);
        }
    };
    int qs, rs;
    std::vector<uint64_t> qf, rf;
    std::vector<int32_t> qc, rc2;
    layout(nq, q_hashes, q_offsets, qs, qf, qc);
    layout(nr, r_hashes, r_offsets, rs, rf, rc2);
    DevBuf dq, dqc, dr, drc;
    struct Free {
        DevBuf *b[4];
        ~Free()
        {
            for (auto *x : b) x->release();
        }
    } fr{{&dq, &dqc, &dr, &drc}};
    HCHECK(dq.ensure(qf.size() * 8), "alloc");
    HCHECK(dqc.ensure(qc.size() * 4), "alloc");
    HCHECK(dr.ensure(rf.size() * 8), "alloc");
    HCHECK(drc.ensure(rc2.size() * 4), "alloc");
    HCHECK(hipMemcpy(dq.p, qf.data(), qf.size() * 8, hipMemcpyHostToDevice), "upload");
    HCHECK(hipMemcpy(dqc.p, qc.data(), qc.size() * 4, hipMemcpyHostToDevice), "upload");
    HCHECK(hipMemcpy(dr.p, rf.data(), rf.size() * 8, hipMemcpyHostToDevice), "upload");
    HCHECK(hipMemcpy(drc.p, rc2.data(), rc2.size() * 4, hipMemcpyHostToDevice), "upload");
    std::vector<int32_t> b, num, den;
    const int r = mash_select_device(ctx, nq, dq, dqc, qs, nr, dr, drc, rs, b, num, den);
    if (r != HALIGN_OK) return r;
    for (int i = 0; i < nq; ++i) {
        best[i] = b[i];
        if (jaccard) jaccard[i] = (double)num[i] / (double)den[i];
    }
    return HALIGN_OK;
}

}  // extern "C"
