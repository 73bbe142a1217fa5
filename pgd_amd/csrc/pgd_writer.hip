// pgd_writer.hip -- the .micro text on the device (write_micro_stats and
// print_micro_stats, graphlet_core.cpp:1668-1761, 1798-1827).
//
// Each E_ordered position e gives one row
//   e_v+1,e_u+1,tri,2-star,4-clique,4-chordal-cycle,4-tailed-triangle,4-cycle,3-star,4-path\n
// (the first two fields only with output_id), every value in decimal as
// ostream prints an unsigned long long.  The rows of a chunk of positions are
// formatted in three steps:
//   k_row_len   the byte length of each row;
//   a scan      the row offsets within the chunk's text;
//   k_row_text  each 256-row block formats its rows into LDS at their
//               offsets, then stores the block's bytes to HBM as aligned
//               32-bit words (byte stores only where a word is shared with a
//               neighbouring block).
// The host copies each chunk's text to one of two pinned buffers and writes
// it to the caller's file descriptor while the next chunk is formatted and
// copied.
#include <unistd.h>

#include <cerrno>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "pgd_internal.hpp"

namespace pgd {
namespace {

constexpr int kRowsPerBlock = 256;
constexpr int kMaxRowBytes = 10 * 20 + 10;  // ten 20-digit fields, nine commas, a newline
constexpr int64_t kChunkRows = (int64_t)1 << 20;
// the order of the columns in a row (indices into the PGD_MICRO_* columns)
__constant__ int kColOrder[8] = {PGD_MICRO_TRI,    PGD_MICRO_STAR2, PGD_MICRO_CLIQUE, PGD_MICRO_CHORDAL,
                                 PGD_MICRO_TAILED, PGD_MICRO_CYCLE, PGD_MICRO_STAR3,  PGD_MICRO_PATH};

__device__ __forceinline__ int ndigits(uint64_t v) {
    int d = 1;
    uint64_t p = 10;
    while (d < 20 && v >= p) {
        d++;
        p *= 10;
    }
    return d;
}

// the digits of v, most significant first, ending at p + ndigits(v)
__device__ __forceinline__ void put_digits(char* p, int nd, uint64_t v) {
    char* q = p + nd;
    while (v >= (1ull << 32)) {
        const uint64_t t = v / 10;
        *--q = (char)('0' + (int)(v - 10 * t));
        v = t;
    }
    uint32_t w = (uint32_t)v;
    do {
        const uint32_t t = w / 10;
        *--q = (char)('0' + (int)(w - 10 * t));
        w = t;
    } while (w);
}

struct RowSrc {
    const int32_t* order;
    const int32_t* ev;
    const int32_t* eu;
    const uint64_t* micro;  // PGD_NMICRO columns of m values
    int64_t m;
    int output_id;
};

__device__ __forceinline__ uint64_t field(const RowSrc& r, int64_t e, int f) {
    if (r.output_id) {
        if (f == 0) return (uint64_t)r.ev[r.order[e]] + 1;
        if (f == 1) return (uint64_t)r.eu[r.order[e]] + 1;
        f -= 2;
    }
    return r.micro[(size_t)kColOrder[f] * r.m + e];
}

__global__ __launch_bounds__(kRowsPerBlock) void k_row_len(RowSrc r, int64_t e0, int64_t rows,
                                                           uint32_t* __restrict__ len) {
    const int64_t i = blockIdx.x * (int64_t)kRowsPerBlock + threadIdx.x;
    if (i >= rows) return;
    const int nf = r.output_id ? 10 : 8;
    uint32_t n = (uint32_t)nf;  // nf - 1 commas and a newline
    for (int f = 0; f < nf; f++) n += (uint32_t)ndigits(field(r, e0 + i, f));
    len[i] = n;
}

// off: exclusive offsets of the rows of the chunk (rows + 1 values)
__global__ __launch_bounds__(kRowsPerBlock) void k_row_text(RowSrc r, int64_t e0, int64_t rows,
                                                            const uint64_t* __restrict__ off,
                                                            char* __restrict__ out) {
    __shared__ char buf[kRowsPerBlock * kMaxRowBytes];
    const int64_t b0 = blockIdx.x * (int64_t)kRowsPerBlock;
    const int64_t i = b0 + threadIdx.x;
    const int64_t last = min(rows, b0 + kRowsPerBlock);
    const uint64_t base = off[b0];
    const uint64_t len = off[last] - base;
    if (i < rows) {
        char* p = buf + (off[i] - base);
        const int nf = r.output_id ? 10 : 8;
        for (int f = 0; f < nf; f++) {
            const uint64_t v = field(r, e0 + i, f);
            const int nd = ndigits(v);
            put_digits(p, nd, v);
            p += nd;
            *p++ = f + 1 < nf ? ',' : '\n';
        }
    }
    __syncthreads();
    // the block's bytes [base, base + len) of `out`, as the 32-bit words that
    // cover them; a word that the block covers only in part is written byte
    // by byte (its other bytes belong to the neighbouring blocks)
    const uint64_t w0 = base & ~(uint64_t)3, w1 = (base + len + 3) & ~(uint64_t)3;
    for (uint64_t a = w0 + 4 * (uint64_t)threadIdx.x; a < w1; a += 4 * kRowsPerBlock) {
        if (a >= base && a + 4 <= base + len) {
            const char* s = buf + (a - base);
            const uint32_t word = (uint32_t)(uint8_t)s[0] | (uint32_t)(uint8_t)s[1] << 8 |
                                  (uint32_t)(uint8_t)s[2] << 16 | (uint32_t)(uint8_t)s[3] << 24;
            *(uint32_t*)(out + a) = word;
        } else {
            for (int j = 0; j < 4; j++) {
                const uint64_t x = a + (uint64_t)j;
                if (x >= base && x < base + len) out[x] = buf[x - base];
            }
        }
    }
}

__global__ void k_zero_first(uint64_t* off) {
    if (threadIdx.x == 0) off[0] = 0;
}

// every byte of buf to fd, retrying short writes
void write_all(pgd_ctx* ctx, int fd, const char* buf, size_t n) {
    while (n) {
        const ssize_t k = ::write(fd, buf, n);
        if (k < 0 && errno == EINTR) continue;
        PGD_CHECK(k > 0, PGD_ERR_INVALID, std::string("write failed: ") + std::strerror(errno));
        buf += k;
        n -= (size_t)k;
    }
    (void)ctx;
}

}  // namespace

void write_micro_text(pgd_ctx* ctx, const GraphView& g, const uint64_t* micro, int fd, bool output_id,
                      const char* prefix, int64_t* written) {
    hipStream_t s = ctx->stream;
    std::string head = prefix ? prefix : "";
    if (output_id) head += "src,dst,";
    head += "triangle,2-star,4-clique,4-chordal-cycle,4-tailed-triangle,4-cycle,3-star,4-path\n";
    write_all(ctx, fd, head.data(), head.size());
    int64_t total = (int64_t)head.size();
    const int64_t rows = g.nord;
    if (rows > 0) {
        const RowSrc src{g.order, g.ev, g.eu, micro, g.m, output_id ? 1 : 0};
        const int64_t crows = std::min(rows, kChunkRows);
        const size_t cap = (size_t)crows * kMaxRowBytes;
        DBuf<uint32_t> len;
        DBuf<uint64_t> off;
        DBuf<char> text[2];
        len.alloc(ctx, crows);
        off.alloc(ctx, crows + 1);
        text[0].alloc(ctx, cap);
        text[1].alloc(ctx, cap);
        char* host[2] = {nullptr, nullptr};
        hipEvent_t done[2] = {nullptr, nullptr};
        uint64_t* tot = nullptr;  // pinned: the chunk's text length
        auto cleanup = [&] {
            for (int k = 0; k < 2; k++) {
                if (done[k]) (void)hipEventDestroy(done[k]);
                if (host[k]) (void)hipHostFree(host[k]);
            }
            if (tot) (void)hipHostFree(tot);
        };
        try {
            for (int k = 0; k < 2; k++) {
                PGD_HIP(hipHostMalloc((void**)&host[k], cap, hipHostMallocDefault));
                PGD_HIP(hipEventCreateWithFlags(&done[k], hipEventDisableTiming));
            }
            PGD_HIP(hipHostMalloc((void**)&tot, sizeof(uint64_t), hipHostMallocDefault));
            size_t scan_bytes = 0;
            PGD_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, len.get(), off.get() + 1, (size_t)crows,
                                            rocprim::plus<uint64_t>(), s));
            void* scan_tmp = tmp_storage(ctx, scan_bytes);
            size_t pending = 0;  // bytes of the chunk in flight to the other pinned buffer
            int slot = 0;
            for (int64_t e0 = 0; e0 < rows; e0 += crows, slot ^= 1) {
                const int64_t n = std::min(crows, rows - e0);
                const int blocks = (int)((n + kRowsPerBlock - 1) / kRowsPerBlock);
                // host[slot] was written out two chunks ago (below), and its copy finished then
                hipLaunchKernelGGL(k_row_len, dim3(blocks), dim3(kRowsPerBlock), 0, s, src, e0, n, len.get());
                hipLaunchKernelGGL(k_zero_first, dim3(1), dim3(64), 0, s, off.get());
                PGD_HIP(hipGetLastError());
                PGD_HIP(rocprim::inclusive_scan(scan_tmp, scan_bytes, len.get(), off.get() + 1, (size_t)n,
                                                rocprim::plus<uint64_t>(), s));
                PGD_HIP(hipMemcpyAsync(tot, off.get() + n, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
                PGD_HIP(hipStreamSynchronize(s));  // (the previous chunk's copy is ahead of it on s)
                const size_t bytes = (size_t)*tot;
                PGD_CHECK(bytes <= cap, PGD_ERR_HIP, "micro text longer than its bound");
                hipLaunchKernelGGL(k_row_text, dim3(blocks), dim3(kRowsPerBlock), 0, s, src, e0, n, off.get(),
                                   text[slot].get());
                PGD_HIP(hipGetLastError());
                PGD_HIP(hipMemcpyAsync(host[slot], text[slot].get(), bytes, hipMemcpyDeviceToHost, s));
                PGD_HIP(hipEventRecord(done[slot], s));
                // the previous chunk goes to the file while this one is formatted and copied
                if (pending) {
                    PGD_HIP(hipEventSynchronize(done[slot ^ 1]));
                    write_all(ctx, fd, host[slot ^ 1], pending);
                    total += (int64_t)pending;
                }
                pending = bytes;
            }
            PGD_HIP(hipEventSynchronize(done[slot ^ 1]));
            write_all(ctx, fd, host[slot ^ 1], pending);
            total += (int64_t)pending;
        } catch (...) {
            (void)hipStreamSynchronize(s);
            cleanup();
            throw;
        }
        cleanup();
    }
    if (written) *written = total;
}

}  // namespace pgd
