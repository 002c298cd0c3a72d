// toplist.hip -- the bank toplist on the device: Toplist::thresholds and Toplist::retain
// (src/workload.cpp:126-141) applied in the launch stream after each batch, so that the
// candidates of a bank never travel to the host.  The result equals the host toplist
// (brp_toplist_retain) row for row:
//
//   thresholds  p.threshold[h] = max(p.threshold[h], last row of level h) for every slot
//               of a batch, from the last rows a retain recorded (never above the
//               one-template-at-a-time threshold, so candidates are a superset)
//   retain      per template in slot order, per level in bin order: skip a candidate at or
//               below the template's threshold or the level's last row; a bin already on
//               the level is replaced only by a stronger detection; else the last row is;
//               the rewritten row moves once, stably, to after every row of greater power
//               and every row of equal power that stood before it (the rows of glibc's
//               stable qsort, as the host retain)
//
// One 256-thread block per (toplist, level): levels never interact, so the five levels of
// every toplist are retained concurrently.  The level's rows stay in LDS for the batch;
// each template's candidates of the level are gathered and bitonic-sorted by bin in LDS.
// Errors (flags in *error): 1 a slot with an invalid resample index or a candidate
// overflow, 2 an invalid candidate (level >= 5, bin outside [W, F), non-finite power),
// 4 more than kMaxLevelCandidates candidates of one level in one template.
#include "device.h"
#include "engine.h"

namespace {

constexpr int kThreads = 256;

// Block-wide reductions over kThreads threads (four waves): scratch holds one word per wave
// plus the result.
__device__ __forceinline__ uint32_t block_min(uint32_t v, uint32_t* scratch) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, uint32_t(__shfl_xor(int(v), off, 64)));
    const int t = threadIdx.x;
    if ((t & 63) == 0) scratch[t >> 6] = v;
    __syncthreads();
    const uint32_t r = min(min(scratch[0], scratch[1]), min(scratch[2], scratch[3]));
    __syncthreads();
    return r;
}

__device__ __forceinline__ uint32_t block_sum(uint32_t v, uint32_t* scratch) {
    for (int off = 32; off > 0; off >>= 1) v += uint32_t(__shfl_xor(int(v), off, 64));
    const int t = threadIdx.x;
    if ((t & 63) == 0) scratch[t >> 6] = v;
    __syncthreads();
    const uint32_t r = scratch[0] + scratch[1] + scratch[2] + scratch[3];
    __syncthreads();
    return r;
}

// lowest: the last row's power of each (table, level), as a retain left it.
__global__ __launch_bounds__(kThreads) void k_thresholds(brp_params* __restrict__ params, uint32_t count,
                                                         const uint32_t* __restrict__ slot_table,
                                                         const float* __restrict__ lowest) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    const float* low = lowest + size_t(slot_table[i]) * 5;
    brp_params& p = params[i];
#pragma unroll
    for (int h = 0; h < 5; ++h) p.threshold[h] = fmaxf(p.threshold[h], low[h]);
}

constexpr uint32_t kMaxRows = brp_engine::kDeviceToplistMaxRows;
constexpr uint32_t kMaxLevelCandidates = 4096;
constexpr uint32_t kRowsPerThread = kMaxRows / kThreads;

// grid (tables, 5), kThreads threads.
__global__ __launch_bounds__(kThreads) void k_retain(const uint32_t* __restrict__ results, uint32_t slot_words,
                                                     const brp_params* __restrict__ params, uint32_t count,
                                                     const uint32_t* __restrict__ slot_table,
                                                     const float4* __restrict__ slot_orbit,
                                                     brp_engine::TopRow* __restrict__ tables, uint32_t rows,
                                                     float* __restrict__ lowest, uint32_t* __restrict__ error) {
    __shared__ brp_engine::TopRow level[kMaxRows];
    __shared__ uint2 list[kMaxLevelCandidates];  // (bin, power bits)
    __shared__ uint32_t scratch[8];
    __shared__ uint32_t gathered;
    const uint32_t table = blockIdx.x, h = blockIdx.y;
    const int t = threadIdx.x;
    brp_engine::TopRow* dst = tables + (size_t(table) * 5 + h) * rows;
    for (uint32_t r = t; r < rows; r += kThreads) level[r] = dst[r];
    __syncthreads();

    for (uint32_t i = 0; i < count; ++i) {
        if (slot_table[i] != table) continue;
        const uint32_t* rec = results + size_t(i) * slot_words;
        const brp_params& p = params[i];
        const uint32_t n = rec[0], flags = rec[1];
        if (flags || n > p.capacity) {
            if (t == 0 && h == 0) atomicOr(error, 1u);
            continue;
        }
        // Gather this level's candidates (and check every candidate once, on level 0).
        if (t == 0) gathered = 0;
        __syncthreads();
        for (uint32_t c = t; c < n; c += kThreads) {
            const uint4 cand = reinterpret_cast<const uint4*>(rec + 4)[c];
            const float power = __uint_as_float(cand.z);
            if (h == 0 && (cand.y >= 5 || cand.x < p.window || cand.x >= p.fundamental || !isfinite(power)))
                atomicOr(error, 2u);
            if (cand.y == h && cand.x >= p.window && cand.x < p.fundamental && isfinite(power)) {
                const uint32_t at = atomicAdd(&gathered, 1u);
                if (at < kMaxLevelCandidates) list[at] = make_uint2(cand.x, cand.z);
            }
        }
        __syncthreads();
        const uint32_t m = gathered;
        __syncthreads();  // every thread has read the count before the next template resets it
        if (m > kMaxLevelCandidates) {
            if (t == 0) atomicOr(error, 4u);
            continue;
        }
        if (m == 0) continue;
        // Bitonic sort by bin (bins of one level and template are distinct), padded to a
        // power of two with bins that sort last.
        uint32_t len = 1;
        while (len < m) len <<= 1;
        for (uint32_t k = m + t; k < len; k += kThreads) list[k] = make_uint2(0xFFFFFFFFu, 0u);
        __syncthreads();
        for (uint32_t size = 2; size <= len; size <<= 1)
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                for (uint32_t k = t; k < len; k += kThreads) {
                    const uint32_t partner = k ^ stride;
                    if (partner > k) {
                        const bool up = (k & size) == 0;
                        const uint2 a = list[k], b = list[partner];
                        if ((a.x > b.x) == up) { list[k] = b; list[partner] = a; }
                    }
                }
                __syncthreads();
            }
        const float thr = p.threshold[h];
        const float4 orbit = slot_orbit[i];
        for (uint32_t k = 0; k < m; ++k) {
            const uint32_t bin = list[k].x;
            const float power = __uint_as_float(list[k].y);
            if (power <= thr || power <= level[rows - 1].power) continue;  // uniform
            // A bin already on the level: its first row; replaced only by a stronger one.
            uint32_t found = rows;
            for (uint32_t r = t; r < rows; r += kThreads)
                if (level[r].bin == bin) found = min(found, r);
            found = block_min(found, scratch);
            uint32_t place = rows - 1;
            if (found < rows) {
                if (!(level[found].power < power)) continue;
                place = found;
            }
            // Insertion point in the level with row `place` removed (closed[k] = level[k]
            // below place, level[k + 1] from place on): the rows of greater power and
            // those of equal power that stood before the rewritten row.
            uint32_t before = 0;
            for (uint32_t r = t; r + 1 < rows; r += kThreads) {
                const float q = level[r < place ? r : r + 1].power;
                before += (q > power || (q == power && r < place)) ? 1u : 0u;
            }
            const uint32_t at = block_sum(before, scratch);
            // New level: closed[0 .. at), the new row, closed[at .. rows - 1).
            brp_engine::TopRow next[kRowsPerThread];
#pragma unroll
            for (uint32_t u = 0; u < kRowsPerThread; ++u) {
                const uint32_t r = t + u * kThreads;
                if (r >= rows) break;
                if (r == at) {
                    next[u] = brp_engine::TopRow{power, orbit.x, orbit.y, orbit.z, bin, 1u << h};
                } else {
                    const uint32_t c = r < at ? r : r - 1;  // index in the closed level
                    next[u] = level[c < place ? c : c + 1];
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t u = 0; u < kRowsPerThread; ++u) {
                const uint32_t r = t + u * kThreads;
                if (r >= rows) break;
                level[r] = next[u];
            }
            __syncthreads();
        }
    }
    for (uint32_t r = t; r < rows; r += kThreads) dst[r] = level[r];
    if (t == 0) lowest[size_t(table) * 5 + h] = level[rows - 1].power;
}

}  // namespace

namespace brp_engine {

void toplist_thresholds(hipStream_t st, brp_params* params, uint32_t count, const uint32_t* slot_table,
                        const float* lowest) {
    if (!count) return;
    k_thresholds<<<(count + kThreads - 1) / kThreads, kThreads, 0, st>>>(params, count, slot_table, lowest);
}

void toplist_retain(hipStream_t st, const uint32_t* results, uint32_t slot_words, const brp_params* params,
                    uint32_t count, const uint32_t* slot_table, const float4* slot_orbit, uint32_t tables,
                    TopRow* rows_dev, uint32_t rows, float* lowest, uint32_t* error) {
    if (!count || !tables) return;
    k_retain<<<dim3(tables, 5), kThreads, 0, st>>>(results, slot_words, params, count, slot_table, slot_orbit,
                                                  rows_dev, rows, lowest, error);
}

}  // namespace brp_engine
