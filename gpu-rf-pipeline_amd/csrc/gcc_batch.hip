// Batched GCC-PHAT over many capture pairs of uneven length, with the peak search fused in.
//
// The reference runs gcc_phat_cuda (src/xcorr.cu:379-399) once per pair and finds the peak on
// the host (src/main.cpp:390-401). Here every pair's transform is cut into work items of the
// same size, 4096 complex values, so that one persistent launch per pass covers the whole
// batch whatever the pair lengths:
//   pass 1  column FFTs of a and b (four-step column pass, C columns x n2 rows per item);
//   pass 2  row FFTs of a and b, PHAT weighting, inverse row FFT, conjugate twiddle
//           (F rows x n1 per item);
//   pass 3  inverse column FFTs; every lag's |r| goes into a per-pair atomic max of
//           (magnitude, first index) instead of being stored (unless the caller asks for r);
//   finish  per pair: |r| at the peak's two neighbours recomputed from the pass-2 columns, and
//           the sub-sample offset in float64 from the two-sample sinc ratio: a whitened pure
//           delay d has |r[k]| = |sinc(k - d)|, so with m0 = |r[peak]| and m1 the larger
//           neighbour, |sinc(1 - f)| / |sinc(f)| = f / (1 - f) gives f = m1 / (m0 + m1),
//           towards that neighbour.
// Pairs need 4096 <= L <= 2^24 (L = next power of two >= Na + Nb - 1, so Na + Nb - 1 > 2048);
// GCC-PHAT's weights depend on L, so shorter pairs are not padded up but refused (the
// single-pair rf_xcorr_fft handles them in one workgroup). All items cost the same; a
// persistent grid walks them in XCD-major order. Pairs whose rows are 4096 points (L >= 2^16)
// run their row pass as one 4096-point row per item on 256 lanes (register-resident radix-16
// transforms); shorter rows use the generic LDS transform, F rows per item. Pairs with
// L = 2^15 skip the passes: one workgroup per pair does their whole GCC-PHAT and peak from
// registers and LDS (gcc_pair32k.hpp), reading each input from HBM once per half-spectrum.

#include "gcc_batch_kernels.hpp"
#include "gcc_pair32k.hpp"
#include "twiddle.hpp"

#include <algorithm>
#include <map>
#include <string>
#include <vector>

using rf::PairDesc;

struct rf_gcc_batch {
    int count = 0;
    std::vector<PairDesc> host;
    PairDesc* dev = nullptr;
    int2* col_map = nullptr;    // col_total items: {pair, column tile} of the inverse pass
    int col_total = 0;
    int2* fwd_map = nullptr;    // the forward pass's items (wider tiles)
    // Row-pass items, split by row length: pairs with 4096-point rows run the register-resident
    // radix-16 kernel, the others the generic LDS kernel, each over its own item map.
    struct RowList {
        int count = 0;
        int total = 0;
        int2* map = nullptr;    // total items: {pair, row item within the pair}
    };
    RowList rows4096;
    RowList rows_generic;
    // Host copies of the item prefixes (column items per pair, row items per list entry) and
    // the pair index of each list entry, for cutting the batch into groups.
    std::vector<int> col_prefix_host;
    std::vector<int> fwd_prefix_host;
    std::vector<int> row_prefix_host[2];
    std::vector<int> row_pairs_host[2];
    // The batch runs as consecutive groups of pairs whose work buffers (2 L values per pair)
    // fit a byte budget, so that the strided column passes and the row pass of a group find
    // their intermediate data in the 256 MB Infinity Cache instead of HBM pages. Budget 0:
    // one group.
    struct Group {
        int col_begin, col_items;
        int fwd_begin, fwd_items;
        int row_begin[2], row_items[2];
    };
    std::vector<Group> groups;
    std::size_t group_budget = 0;
    float2* work = nullptr;
    // Pairs with L = 2^15: one workgroup each (gcc_pair32k.hpp).
    int fused_count = 0;
    int* fused_list = nullptr;  // their pair indices
    float2* scratch = nullptr;  // pair32k_scratch_values()
    std::map<int, float2*> tables;
    std::size_t bytes = 0;
    // workgroups per CU of the row and inverse passes (more than fit at once: the surplus
    // queues behind the resident ones, which evens out the uneven items); rf_gcc_batch_tune
    // picks from 2..8, and 8 was the fastest on the 256 uneven pairs it was tuned on
    int grid_mult = 8;
    hipStream_t capture = nullptr;
    hipGraphExec_t graph = nullptr;
    // When the batch has pairs of both kinds, the single-workgroup pairs run on a side stream
    // beside the passes (forked from and joined back into the caller's stream), so that their
    // one-pair-per-CU kernel fills in where the passes leave room instead of adding its whole
    // length after them.
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr;
    hipEvent_t join = nullptr;
    rf_peak_result* graph_results = nullptr;
};

namespace rf {
namespace {

__global__ __launch_bounds__(kItemThreads) void batch_row_kernel(
    const PairDesc* __restrict__ pairs, const int2* __restrict__ map, int begin, int total,
    const float2* __restrict__ table)
{
    extern __shared__ float2 lds[];
    xcd_major_items(total, [&](int local) {
        const int2 m = map[begin + local];
        const PairDesc d = pairs[m.x];
        const FourStep fs = shape_of(d);
        const long long L = 1LL << d.log2L;
        const int n1 = 1 << d.log2n1;
        const int F = lds_batch_per_block(d.log2n1);
        const int slots = lds_slots(n1);
        const int row0 = m.y * F;
        float2* wa = d.work + static_cast<long long>(row0) * n1;
        float2* wb = d.work + L + static_cast<long long>(row0) * n1;
        float2* la = lds;
        float2* lb = lds + F * slots;
        const int total_e = F * n1;
        for (int e = threadIdx.x; e < total_e; e += blockDim.x) {
            const int s = (e >> d.log2n1) * slots + lds_pad(e & (n1 - 1));
            la[s] = gload(wa, e);
            lb[s] = gload(wb, e);
        }
        __syncthreads();
        const int per = n1 >> 3;
        const int f = threadIdx.x / per;
        const int t = threadIdx.x - f * per;
        lds_fft<false>(la + f * slots, d.log2n1, t, true, table);
        lds_fft<false>(lb + f * slots, d.log2n1, t, true, table);
        const float scale = 1.0f / static_cast<float>(L);
        for (int e = threadIdx.x; e < total_e; e += blockDim.x) {
            const int s = (e >> d.log2n1) * slots + lds_pad(e & (n1 - 1));
            lb[s] = weigh(la[s], lb[s], scale, 1);
        }
        __syncthreads();
        lds_fft<true>(lb + f * slots, d.log2n1, t, true, table);
        for (int e = threadIdx.x; e < total_e; e += blockDim.x) {
            const int rr = e >> d.log2n1;
            const int j1 = e & (n1 - 1);
            const float2 v = lb[rr * slots + lds_pad(j1)];
            gstore(wb, e, cmul(v, four_step_twiddle<true>(fs, static_cast<long long>(j1) * (row0 + rr))));
        }
        __syncthreads();
    });
}

// Row pass of the pairs whose rows are 4096 points: one row per item, 256 lanes, the
// register-resident radix-16 transforms. Built for three waves per SIMD (at most 168
// registers, which it fits without spilling) instead of the two its unconstrained allocation
// (173) would allow.
__global__ __launch_bounds__(k4096Lanes, 3) void batch_row4096_kernel(
    const PairDesc* __restrict__ pairs, const int2* __restrict__ map, int begin, int total,
    const float2* __restrict__ stages)
{
    __shared__ float2 buf[k4096Slots];
    xcd_major_items(total, [&](int local) {
        const int2 m = map[begin + local];
        const PairDesc d = pairs[m.x];
        const FourStep fs = shape_of(d);
        const long long L = 1LL << d.log2L;
        const long long row = m.y;
        const float2* wa = d.work + (row << k4096Log2);
        float2* wb = d.work + L + (row << k4096Log2);
        xcorr_row4096(wa, wb, fs, row, 1, 1.0f / static_cast<float>(L), buf, stages);
        __syncthreads();  // the next item reuses buf
    });
}

std::size_t row_lds_bytes()
{
    // 2 F lds_slots(n1) with F n1 = 4096: at most 2 (4096 + 512 + F), F <= 4.
    return static_cast<std::size_t>(2 * (4096 + 512 + 4)) * sizeof(float2);
}

constexpr std::size_t kDefaultGroupBytes = 0;  // one group: measured fastest on MI355X

void make_groups(rf_gcc_batch& h, std::size_t budget)
{
    h.group_budget = budget;
    h.groups.clear();
    int p0 = 0;
    while (p0 < h.count) {
        int p1 = p0;
        std::size_t bytes = 0;
        while (p1 < h.count) {
            const PairDesc& d = h.host[static_cast<std::size_t>(p1)];
            const std::size_t pair_bytes =
                d.fused != 0 ? 0 : (static_cast<std::size_t>(2) << d.log2L) * sizeof(float2);
            if (p1 > p0 && budget != 0 && bytes + pair_bytes > budget) {
                break;
            }
            bytes += pair_bytes;
            ++p1;
        }
        rf_gcc_batch::Group g{};
        g.col_begin = h.col_prefix_host[static_cast<std::size_t>(p0)];
        g.col_items = h.col_prefix_host[static_cast<std::size_t>(p1)] - g.col_begin;
        g.fwd_begin = h.fwd_prefix_host[static_cast<std::size_t>(p0)];
        g.fwd_items = h.fwd_prefix_host[static_cast<std::size_t>(p1)] - g.fwd_begin;
        for (int kind = 0; kind < 2; ++kind) {
            const std::vector<int>& pairs = h.row_pairs_host[kind];
            const std::vector<int>& prefix = h.row_prefix_host[kind];
            const auto lo = std::lower_bound(pairs.begin(), pairs.end(), p0) - pairs.begin();
            const auto hi = std::lower_bound(pairs.begin(), pairs.end(), p1) - pairs.begin();
            g.row_begin[kind] = prefix[static_cast<std::size_t>(lo)];
            g.row_items[kind] = prefix[static_cast<std::size_t>(hi)] - g.row_begin[kind];
        }
        h.groups.push_back(g);
        p0 = p1;
    }
}

void batch_run(rf_gcc_batch& h, rf_peak_result* results, hipStream_t stream)
{
    const float2* table = twiddle_table_4096();
    const int cus = cu_count();
    batch_init_kernel<<<static_cast<unsigned>(ceil_div(h.count, 256)), 256, 0, stream>>>(results,
                                                                                         h.count);
    check_launch("batch_init_kernel");
    const bool forked = h.fused_count > 0 && h.side != nullptr;
    if (forked) {
        RF_HIP_CHECK(hipEventRecord(h.fork, stream));
        RF_HIP_CHECK(hipStreamWaitEvent(h.side, h.fork, 0));
    }
    if (h.fused_count > 0) {
        launch_pair32k(h.dev, h.fused_list, static_cast<const int*>(nullptr), h.fused_count,
                       h.fused_count, results, h.scratch, forked ? h.side : stream);
    }
    if (forked) {
        RF_HIP_CHECK(hipEventRecord(h.join, h.side));
    }
    // Persistent grids: grid_mult workgroups per CU, never more than the items, rounded up to a
    // multiple of 8 for the XCD-major walk (surplus workgroups find no item).
    const int cap = h.grid_mult * cus;
    auto grid_for = [cap](int items) {
        return static_cast<unsigned>(ceil_div(std::min(items, cap), 8) * 8);
    };
    // the forward column pass's 80 KiB workgroups: never more than fit on the CUs at once
    const int fwd_cap = std::min(h.grid_mult, kFwdWorkgroupsPerCu) * cus;
    auto grid_fwd = [fwd_cap](int items) {
        return static_cast<unsigned>(ceil_div(std::min(items, fwd_cap), 8) * 8);
    };
    allow_large_lds(reinterpret_cast<const void*>(batch_column_fwd_kernel),
                    column_fwd_lds_bytes());
    for (const rf_gcc_batch::Group& g : h.groups) {
        if (g.col_items == 0) {
            continue;  // only pairs of the single-workgroup path
        }
        batch_column_fwd_kernel<<<grid_fwd(2 * g.fwd_items), kFwdThreads,
                                  column_fwd_lds_bytes(), stream>>>(h.dev, h.fwd_map, g.fwd_begin,
                                                                    g.fwd_items, table);
        check_launch("batch_column_fwd_kernel");
        if (g.row_items[0] > 0) {
            batch_row4096_kernel<<<grid_for(g.row_items[0]), k4096Lanes, 0, stream>>>(
                h.dev, h.rows4096.map, g.row_begin[0], g.row_items[0], fft4096_stage_table());
            check_launch("batch_row4096_kernel");
        }
        if (g.row_items[1] > 0) {
            batch_row_kernel<<<grid_for(g.row_items[1]), kItemThreads, row_lds_bytes(), stream>>>(
                h.dev, h.rows_generic.map, g.row_begin[1], g.row_items[1], table);
            check_launch("batch_row_kernel");
        }
        batch_column_inv_peak_kernel<<<grid_for(g.col_items), kItemThreads,
                                       column_lds_bytes(), stream>>>(h.dev, h.col_map, g.col_begin, g.col_items,
                                                 results, ItemSet{nullptr, nullptr, 0}, table);
        check_launch("batch_column_inv_peak_kernel");
    }
    batch_finish_kernel<<<static_cast<unsigned>(h.count), 64, 0, stream>>>(h.dev, results,
                                                                         nullptr);
    check_launch("batch_finish_kernel");
    if (forked) {
        RF_HIP_CHECK(hipStreamWaitEvent(stream, h.join, 0));
    }
}

void destroy_batch(rf_gcc_batch* h)
{
    if (h == nullptr) {
        return;
    }
    if (h->graph != nullptr) {
        (void)hipGraphExecDestroy(h->graph);
    }
    if (h->capture != nullptr) {
        (void)hipStreamDestroy(h->capture);
    }
    if (h->side != nullptr) {
        (void)hipStreamDestroy(h->side);
    }
    if (h->fork != nullptr) {
        (void)hipEventDestroy(h->fork);
    }
    if (h->join != nullptr) {
        (void)hipEventDestroy(h->join);
    }
    (void)hipFree(h->dev);
    (void)hipFree(h->col_map);
    (void)hipFree(h->fwd_map);
    for (auto* list : {&h->rows4096, &h->rows_generic}) {
        (void)hipFree(list->map);
    }
    (void)hipFree(h->work);
    (void)hipFree(h->fused_list);
    (void)hipFree(h->scratch);
    for (auto& kv : h->tables) {
        (void)hipFree(kv.second);
    }
    delete h;
}

}  // namespace
}  // namespace rf

using rf::guarded;

extern "C" {

rf_status rf_gcc_batch_create(int count, const int* num_a, const int* num_b,
                              rf_gcc_batch** out)
{
    return guarded([&] {
        if (out == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: NULL output handle");
        *out = nullptr;
        if (count < 1 || num_a == nullptr || num_b == nullptr) {
            rf::fail(RF_INVALID_ARG, "gcc batch: need at least one pair and its lengths");
        }
        rf_gcc_batch* h = new rf_gcc_batch;
        try {
            h->count = count;
            h->host.resize(static_cast<std::size_t>(count));
            std::vector<int> colp(static_cast<std::size_t>(count) + 1, 0);
            std::vector<int> fwdp(static_cast<std::size_t>(count) + 1, 0);
            std::vector<int> row_pairs[2];            // 4096-point rows, other rows
            std::vector<int> row_prefix[2] = {{0}, {0}};
            std::vector<int> fused;
            long long work_total = 0;
            for (int i = 0; i < count; ++i) {
                if (num_a[i] < 1 || num_b[i] < 1) {
                    rf::fail(RF_INVALID_ARG, "gcc batch: pair " + std::to_string(i)
                                                 + " needs at least one sample in each input");
                }
                const long long lags = static_cast<long long>(num_a[i]) + num_b[i] - 1;
                long long L = 1;
                while (L < lags) {
                    L *= 2;
                }
                if (L < 4096) {
                    rf::fail(RF_LIMIT, "gcc batch: pair " + std::to_string(i) + " has "
                                           + std::to_string(lags) + " lags; batched pairs need "
                                           "more than 2048 (use rf_xcorr_fft)");
                }
                if (L > (1LL << 24)) {
                    rf::fail(RF_LIMIT, "gcc batch: pair " + std::to_string(i) + " needs an FFT "
                                           "longer than the supported 2^24 points");
                }
                const int log2L = rf::ilog2(L);
                const rf::FourStep fs = rf::make_four_step_shape(log2L);
                rf::check_column_tile(fs);
                PairDesc& d = h->host[static_cast<std::size_t>(i)];
                d = PairDesc{};
                d.num_a = num_a[i];
                d.num_b = num_b[i];
                d.log2L = log2L;
                d.log2n1 = fs.log2n1;
                d.log2n2 = fs.log2n2;
                if (log2L == rf::kPair32kLog2L) {
                    d.fused = 1;
                    fused.push_back(i);
                    colp[static_cast<std::size_t>(i) + 1] = colp[static_cast<std::size_t>(i)];
                    fwdp[static_cast<std::size_t>(i) + 1] = fwdp[static_cast<std::size_t>(i)];
                    continue;
                }
                d.col_items = (1 << fs.log2n1) / rf::column_tile(fs);
                d.fwd_items = (1 << fs.log2n1) / rf::column_tile<rf::kFwdTileLog2>(fs);
                fwdp[static_cast<std::size_t>(i) + 1] = fwdp[static_cast<std::size_t>(i)] + d.fwd_items;
                d.row_items = (1 << fs.log2n2) / rf::lds_batch_per_block(fs.log2n1);
                colp[static_cast<std::size_t>(i) + 1] = colp[static_cast<std::size_t>(i)] + d.col_items;
                const int kind = fs.log2n1 == rf::k4096Log2 ? 0 : 1;
                row_pairs[kind].push_back(i);
                row_prefix[kind].push_back(row_prefix[kind].back() + d.row_items);
                work_total += 2 * L;
            }
            h->col_total = colp.back();
            h->col_prefix_host = colp;
            h->fwd_prefix_host = fwdp;
            for (int kind = 0; kind < 2; ++kind) {
                h->row_prefix_host[kind] = row_prefix[kind];
                h->row_pairs_host[kind] = row_pairs[kind];
            }
            rf::make_groups(*h, rf::kDefaultGroupBytes);
            (void)rf::twiddle_table_4096();
            (void)rf::fft4096_stage_table();
            if (work_total > 0) {
                RF_HIP_CHECK(hipMalloc(&h->work, static_cast<std::size_t>(work_total) * sizeof(float2)));
            }
            h->fused_count = static_cast<int>(fused.size());
            if (!fused.empty()) {
                (void)rf::fft16384_table();
                RF_HIP_CHECK(hipMalloc(&h->fused_list, fused.size() * sizeof(int)));
                RF_HIP_CHECK(hipMemcpy(h->fused_list, fused.data(), fused.size() * sizeof(int),
                                       hipMemcpyHostToDevice));
                const std::size_t scratch_bytes =
                    static_cast<std::size_t>(rf::pair32k_scratch_values()) * sizeof(float2);
                RF_HIP_CHECK(hipMalloc(&h->scratch, scratch_bytes));
                h->bytes += scratch_bytes + fused.size() * sizeof(int);
                if (work_total > 0) {
                    RF_HIP_CHECK(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
                    RF_HIP_CHECK(hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
                    RF_HIP_CHECK(hipEventCreateWithFlags(&h->join, hipEventDisableTiming));
                }
            }
            long long offset = 0;
            for (auto& d : h->host) {
                if (d.fused != 0) {
                    continue;
                }
                d.work = h->work + offset;
                offset += 2LL << d.log2L;
                auto it = h->tables.find(d.log2L);
                if (it == h->tables.end()) {
                    float2* t = nullptr;
                    const long long tsize = rf::four_step_table_size(d.log2L);
                    RF_HIP_CHECK(hipMalloc(&t, static_cast<std::size_t>(tsize) * sizeof(float2)));
                    rf::upload_four_step_tables(d.log2L, t);
                    it = h->tables.emplace(d.log2L, t).first;
                    h->bytes += static_cast<std::size_t>(tsize) * sizeof(float2);
                }
                d.lo = it->second;
                d.hi = it->second + rf::kTwiddleSize;
            }
            RF_HIP_CHECK(hipMalloc(&h->dev, h->host.size() * sizeof(PairDesc)));
            auto upload_map = [&](const std::vector<int2>& map, int2** dev) {
                if (map.empty()) {
                    return;  // every pair takes the single-workgroup path
                }
                RF_HIP_CHECK(hipMalloc(dev, map.size() * sizeof(int2)));
                RF_HIP_CHECK(hipMemcpy(*dev, map.data(), map.size() * sizeof(int2),
                                       hipMemcpyHostToDevice));
                h->bytes += map.size() * sizeof(int2);
            };
            std::vector<int2> col_map;
            for (int i = 0; i < count; ++i) {
                for (int c = 0; c < h->host[static_cast<std::size_t>(i)].col_items; ++c) {
                    col_map.push_back(make_int2(i, c));
                }
            }
            upload_map(col_map, &h->col_map);
            std::vector<int2> fwd_map;
            for (int i = 0; i < count; ++i) {
                for (int c = 0; c < h->host[static_cast<std::size_t>(i)].fwd_items; ++c) {
                    fwd_map.push_back(make_int2(i, c));
                }
            }
            upload_map(fwd_map, &h->fwd_map);
            rf_gcc_batch::RowList* lists[2] = {&h->rows4096, &h->rows_generic};
            for (int kind = 0; kind < 2; ++kind) {
                rf_gcc_batch::RowList& list = *lists[kind];
                list.count = static_cast<int>(row_pairs[kind].size());
                list.total = row_prefix[kind].back();
                if (list.count == 0) {
                    continue;
                }
                std::vector<int2> map;
                for (int p : row_pairs[kind]) {
                    for (int r = 0; r < h->host[static_cast<std::size_t>(p)].row_items; ++r) {
                        map.push_back(make_int2(p, r));
                    }
                }
                upload_map(map, &list.map);
            }
            RF_HIP_CHECK(hipMemcpy(h->dev, h->host.data(), h->host.size() * sizeof(PairDesc),
                                   hipMemcpyHostToDevice));
            h->bytes += static_cast<std::size_t>(work_total) * sizeof(float2)
                        + h->host.size() * sizeof(PairDesc);
        } catch (...) {
            rf::destroy_batch(h);
            throw;
        }
        *out = h;
    });
}

rf_status rf_gcc_batch_bind(rf_gcc_batch* h, const rf_cf32* const* a, const rf_cf32* const* b,
                            rf_cf32* const* r)
{
    return guarded([&] {
        if (h == nullptr || a == nullptr || b == nullptr) {
            rf::fail(RF_INVALID_ARG, "gcc batch: NULL handle or channel arrays");
        }
        for (int i = 0; i < h->count; ++i) {
            PairDesc& d = h->host[static_cast<std::size_t>(i)];
            if (a[i] == nullptr || b[i] == nullptr) {
                rf::fail(RF_INVALID_ARG, "gcc batch: NULL input of pair " + std::to_string(i));
            }
            d.a = reinterpret_cast<const float2*>(a[i]);
            d.b = reinterpret_cast<const float2*>(b[i]);
            d.r = r != nullptr ? reinterpret_cast<float2*>(r[i]) : nullptr;
        }
        RF_HIP_CHECK(hipDeviceSynchronize());  // no run may still read the old descriptors
        RF_HIP_CHECK(hipMemcpy(h->dev, h->host.data(), h->host.size() * sizeof(PairDesc),
                               hipMemcpyHostToDevice));
    });
}

rf_status rf_gcc_batch_run(rf_gcc_batch* h, rf_peak_result* results, rf_stream stream)
{
    return guarded([&] {
        if (h == nullptr || results == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: NULL");
        if (h->host[0].a == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: run before bind");
        rf::batch_run(*h, results, rf::as_stream(stream));
    });
}

rf_status rf_gcc_batch_launch_graph(rf_gcc_batch* h, rf_peak_result* results, rf_stream stream)
{
    return guarded([&] {
        if (h == nullptr || results == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: NULL");
        if (h->host[0].a == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: run before bind");
        if (h->graph == nullptr || h->graph_results != results) {
            if (h->graph != nullptr) {
                RF_HIP_CHECK(hipGraphExecDestroy(h->graph));
                h->graph = nullptr;
            }
            if (h->capture == nullptr) {
                RF_HIP_CHECK(hipStreamCreateWithFlags(&h->capture, hipStreamNonBlocking));
            }
            (void)rf::twiddle_table_4096();
            (void)rf::fft4096_stage_table();
            (void)rf::fft16384_table();
            hipGraph_t graph = nullptr;
            RF_HIP_CHECK(hipStreamBeginCapture(h->capture, hipStreamCaptureModeThreadLocal));
            try {
                rf::batch_run(*h, results, h->capture);
            } catch (...) {
                (void)hipStreamEndCapture(h->capture, &graph);
                if (graph != nullptr) (void)hipGraphDestroy(graph);
                throw;
            }
            RF_HIP_CHECK(hipStreamEndCapture(h->capture, &graph));
            const hipError_t err = hipGraphInstantiate(&h->graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            RF_HIP_CHECK(err);
            h->graph_results = results;
        }
        RF_HIP_CHECK(hipGraphLaunch(h->graph, rf::as_stream(stream)));
    });
}

rf_status rf_gcc_batch_tune(rf_gcc_batch* h, rf_peak_result* results, rf_stream stream,
                            int* workgroups_per_cu)
{
    return guarded([&] {
        if (h == nullptr || results == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: NULL");
        if (h->host[0].a == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: tune before bind");
        hipStream_t s = rf::as_stream(stream);
        hipEvent_t start = nullptr;
        hipEvent_t stop = nullptr;
        RF_HIP_CHECK(hipEventCreate(&start));
        RF_HIP_CHECK(hipEventCreate(&stop));
        const int grids[] = {2, 3, 4, 6, 8};
        const std::size_t budgets[] = {0, std::size_t(48) << 20, std::size_t(96) << 20,
                                       std::size_t(160) << 20};
        float best_ms = 0.0f;
        int best = h->grid_mult;
        std::size_t best_budget = h->group_budget;
        try {
            for (std::size_t budget : budgets) {
                rf::make_groups(*h, budget);
                for (int c : grids) {
                    h->grid_mult = c;
                    rf::batch_run(*h, results, s);  // warm-up
                    RF_HIP_CHECK(hipEventRecord(start, s));
                    for (int rep = 0; rep < 3; ++rep) {
                        rf::batch_run(*h, results, s);
                    }
                    RF_HIP_CHECK(hipEventRecord(stop, s));
                    RF_HIP_CHECK(hipEventSynchronize(stop));
                    float ms = 0.0f;
                    RF_HIP_CHECK(hipEventElapsedTime(&ms, start, stop));
                    if (best_ms == 0.0f || ms < best_ms) {
                        best_ms = ms;
                        best = c;
                        best_budget = budget;
                    }
                }
            }
        } catch (...) {
            (void)hipEventDestroy(start);
            (void)hipEventDestroy(stop);
            throw;
        }
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
        rf::make_groups(*h, best_budget);
        h->grid_mult = best;
        if (h->graph != nullptr) {  // the captured graph used the old grid
            RF_HIP_CHECK(hipGraphExecDestroy(h->graph));
            h->graph = nullptr;
        }
        if (workgroups_per_cu != nullptr) {
            *workgroups_per_cu = best;
        }
    });
}

rf_status rf_gcc_batch_set_group_bytes(rf_gcc_batch* h, size_t bytes)
{
    return guarded([&] {
        if (h == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: NULL");
        rf::make_groups(*h, bytes);
        if (h->graph != nullptr) {  // the captured graph used the old groups
            RF_HIP_CHECK(hipGraphExecDestroy(h->graph));
            h->graph = nullptr;
        }
    });
}

rf_status rf_gcc_batch_group_bytes(const rf_gcc_batch* h, size_t* bytes, int* groups)
{
    return guarded([&] {
        if (h == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: NULL");
        if (bytes != nullptr) {
            *bytes = h->group_budget;
        }
        if (groups != nullptr) {
            *groups = static_cast<int>(h->groups.size());
        }
    });
}

rf_status rf_gcc_batch_bytes(const rf_gcc_batch* h, size_t* bytes)
{
    return guarded([&] {
        if (h == nullptr || bytes == nullptr) rf::fail(RF_INVALID_ARG, "gcc batch: NULL");
        *bytes = h->bytes;
    });
}

rf_status rf_gcc_batch_destroy(rf_gcc_batch* h)
{
    return guarded([&] { rf::destroy_batch(h); });
}

}  // extern "C"
