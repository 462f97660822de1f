// Balanced work plan for the z-marching kernel of one domain.
//
// A step's cost is not uniform over the domain: a wave whose sources include an interface cell
// runs the longer interface form (colour gradient, Omega2, recolouring), measured at 1.7x (D3Q19)
// to 2x (D3Q27) the bulk form per plane of a column tile. With uniform z-chunks, a work-group
// whose chunk lies along the interface takes that much longer than one in the bulk, and at sizes
// where every CU holds one work-group (128^3: 256 work-groups) the step waits for the slowest. The plan cuts each column
// tile's z range into chunks of about equal modelled cost instead, from the per-wave interface
// map the kernel writes (MarchArgs::wmap), so that all work-groups of a round finish together.
// The partition only moves work between work-groups: every cell is updated by the same
// arithmetic whatever the chunk bounds, so results are bitwise unchanged.
//
// Cost model, in units of one bulk plane of a column tile: a plane costs 1 + alpha * m / per,
// m = the largest number of interface waves on one SIMD (own wave w runs on SIMD w % 4, per =
// waves per SIMD); a chunk costs its planes plus e (the two edge source planes and the pipeline
// prologue). When the recording step ran on a plan, it also timed that plan's work items, and
// each item's planes are scaled by its measured over modelled cost (relative to all items), so
// what the model misses at a place is corrected at the next cut; the first plan of a state is
// re-made one step later for that. The smallest bound T such that cutting every column greedily into chunks of cost
// <= T needs at most nitems chunks is searched for by trying many bounds at once (one group of
// threads per bound); greedy cutting gives the fewest chunks for a bound, so T is the best bound
// for contiguous chunks.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace cglbm {
namespace {

constexpr int kPlanThreads = 1024;
constexpr int kMaxCand = 32;   // bounds tried at once (one per group of ncol threads)
// Each round narrows the bound interval (ncand + 1)-fold and costs a greedy pass over the whole
// domain on all 16 waves of the planner; rounds stop once T is known to kTolerance planes.
constexpr int kMaxRounds = 16;
constexpr float kTolerance = 0.1f;

// Plane costs of one column in 1/32 of a bulk plane, one byte each, four per 32-bit word.
constexpr float kCostUnit = 1.0f / 32.0f;
__device__ __forceinline__ float plane_cost(uint32_t word, int k) {
    return kCostUnit * static_cast<float>((word >> (8 * k)) & 0xFFu);
}
__device__ __forceinline__ uint32_t cost_byte(float w) {
    const float q = rintf(w * 32.0f);
    return q < 1.0f ? 1u : (q > 255.0f ? 255u : static_cast<uint32_t>(q));
}

// Chunks needed to cut one column greedily so that no chunk costs more than T (the next word of
// costs is loaded while the current one is cut).
__device__ int greedy_count(const uint32_t *col, int nz, float e, float T) {
    int count = 1;
    float acc = e;
    uint32_t next = col[0];
    for (int z4 = 0; z4 < nz; z4 += 4) {
        const uint32_t word = next;
        if (z4 + 4 < nz) next = col[(z4 >> 2) + 1];
        const int kend = nz - z4 < 4 ? nz - z4 : 4;
        for (int k = 0; k < kend; ++k) {
            const float w = plane_cost(word, k);
            if (acc + w > T) {
                ++count;
                acc = e + w;
            } else {
                acc += w;
            }
        }
    }
    return count;
}

// Level of one (column tile, plane) from its NW interface bytes (0/1 per own wave, wave w on
// SIMD w % 4): the most interface waves on one SIMD.
template <int NW>
__device__ __forceinline__ uint32_t plane_level(const uint8_t *w) {
    if constexpr (NW == 8) {
        const uint64_t v = *reinterpret_cast<const uint64_t *>(w);
        const uint32_t s = static_cast<uint32_t>(v) + static_cast<uint32_t>(v >> 32);  // bytes 0..2, no carry
        const uint32_t a = (s & 0xFFu) > ((s >> 8) & 0xFFu) ? (s & 0xFFu) : ((s >> 8) & 0xFFu);
        const uint32_t b = ((s >> 16) & 0xFFu) > (s >> 24) ? ((s >> 16) & 0xFFu) : (s >> 24);
        return a > b ? a : b;
    } else if constexpr (NW == 4) {
        return *reinterpret_cast<const uint32_t *>(w) != 0u ? 1u : 0u;
    } else {
        return *reinterpret_cast<const uint16_t *>(w) != 0u ? 1u : 0u;
    }
}

template <int NW>
__global__ void __launch_bounds__(kPlanThreads)
    plan_kernel(const uint8_t *__restrict__ wmap, const uint32_t *__restrict__ itime, int ntx, int nty, int nz,
                float alpha, float e, int4 *__restrict__ items, int nitems) {
    extern __shared__ int4 smem[];
    const int ncol = ntx * nty;
    const int nzw = (nz + 3) / 4;                        // 32-bit words of levels per column
    const int cs = nzw | 1;                              // column stride, odd: lanes reading one
                                                         // word of different columns hit distinct banks
    int4 *staged = smem;                                 // kPlanThreads items before ordering
    uint32_t *key = reinterpret_cast<uint32_t *>(staged + kPlanThreads);  // kPlanThreads sort keys
    int *best = reinterpret_cast<int *>(key + kPlanThreads);  // ncol chunk counts at the bound s_hi
    int *cand = best + ncol;                             // ncand x ncol counts of one round
    uint32_t *lvl = reinterpret_cast<uint32_t *>(cand + kPlanThreads);  // ncol x cs words of costs
    __shared__ int s_count[kMaxCand];
    __shared__ float s_lo, s_hi, s_sum_t, s_sum_m;
    __shared__ int s_first, s_total;
    __shared__ unsigned int s_maxc;
    const int tid = threadIdx.x;

    constexpr int per = NW == 8 ? 2 : 1;  // own waves per SIMD
    const float unit = alpha / static_cast<float>(per);
    if (tid == 0) {
        s_maxc = 0u;
        s_sum_t = 0.0f;
        s_sum_m = 0.0f;
    }
    for (int i = tid; i < ncol * nzw; i += blockDim.x) {
        const int c = i / nzw, zw = i - c * nzw, z4 = zw * 4;
        uint32_t word = 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (z4 + k < nz) {
                const uint32_t lv = plane_level<NW>(wmap + (static_cast<size_t>(c) * nz + z4 + k) * NW);
                word |= cost_byte(1.0f + unit * static_cast<float>(lv)) << (8 * k);
            }
        lvl[c * cs + zw] = word;
    }
    __syncthreads();

    // Measured correction (itime: durations of the items of the previous plan, still in `items`,
    // timed by the step that wrote the map): each previous item's planes are scaled by its
    // duration over its modelled cost, relative to the same ratio over all items, clamped to
    // [0.5, 2]. The model's errors that belong to a place (the halo work of interface neighbours,
    // the mix of interface waves within a plane) are so carried into the next cut.
    float *ratio = reinterpret_cast<float *>(cand);  // per previous item, before the search uses cand
    if (itime != nullptr) {
        for (int i = tid; i < nitems; i += blockDim.x) {
            const int4 it = items[i];
            float r = 0.0f;
            if (it.z < it.w && it.x >= 0 && it.x < ntx && it.y >= 0 && it.y < nty && it.z >= 0 && it.w <= nz) {
                const uint32_t *col = lvl + (it.y * ntx + it.x) * cs;
                float m = e;
                for (int z = it.z; z < it.w; ++z) m += plane_cost(col[z >> 2], z & 3);
                const float t = static_cast<float>(itime[i]);
                if (t > 0.0f) {
                    r = t / m;
                    atomicAdd(&s_sum_t, t);
                    atomicAdd(&s_sum_m, m);
                }
            }
            ratio[i] = r;
        }
        __syncthreads();
        const float mean = s_sum_m > 0.0f ? s_sum_t / s_sum_m : 0.0f;
        for (int i = tid; i < nitems; i += blockDim.x) {
            const int4 it = items[i];
            if (ratio[i] > 0.0f && mean > 0.0f) {
                float r = ratio[i] / mean;
                r = r < 0.5f ? 0.5f : (r > 2.0f ? 2.0f : r);
                uint32_t *col = lvl + (it.y * ntx + it.x) * cs;
                // planes of one item are this thread's alone (items partition the domain), but a
                // word may hold planes of two items: update its bytes with atomics
                for (int z = it.z; z < it.w; ++z) {
                    const int sh = 8 * (z & 3);
                    const uint32_t old = (col[z >> 2] >> sh) & 0xFFu;
                    const uint32_t nw = cost_byte(static_cast<float>(old) * kCostUnit * r);
                    atomicXor(&col[z >> 2], (old ^ nw) << sh);
                }
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < ncol * nzw; i += blockDim.x) {
        const uint32_t w = lvl[(i / nzw) * cs + i % nzw];
        uint32_t mx = w & 0xFFu;
        mx = ((w >> 8) & 0xFFu) > mx ? ((w >> 8) & 0xFFu) : mx;
        mx = ((w >> 16) & 0xFFu) > mx ? ((w >> 16) & 0xFFu) : mx;
        mx = (w >> 24) > mx ? (w >> 24) : mx;
        atomicMax(&s_maxc, mx);
    }
    __syncthreads();

    // bounds: a chunk of one cheapest plane (lo) and a column of costliest planes (hi, widened
    // against rounding: one chunk per column then fits, feasible as ncol <= nitems)
    for (int c = tid; c < ncol; c += blockDim.x) best[c] = 1;
    if (tid == 0) {
        s_lo = e + kCostUnit;
        s_hi = (e + static_cast<float>(nz) * kCostUnit * static_cast<float>(s_maxc)) * 1.001f;
    }

    // search for the smallest feasible bound: ncand bounds per round, spread over (lo, hi)
    int ncand = static_cast<int>(blockDim.x) / ncol;
    ncand = ncand < 1 ? 1 : (ncand > kMaxCand ? kMaxCand : ncand);
    const int j = tid / ncol;
    for (int round = 0; round < kMaxRounds; ++round) {
        if (tid < kMaxCand) s_count[tid] = 0;
        __syncthreads();
        const float lo = s_lo, hi = s_hi;
        if (hi - lo <= kTolerance) break;  // uniform: every thread read the same s_lo, s_hi
        const float step = (hi - lo) / static_cast<float>(ncand + 1);
        if (j < ncand) {
            const float T = lo + step * static_cast<float>(j + 1);
            int local = 0;
            if (ncand > 1) {
                const int c = tid - j * ncol;
                local = greedy_count(lvl + c * cs, nz, e, T);
                cand[tid] = local;
            } else {
                for (int c = tid; c < ncol; c += blockDim.x) local += greedy_count(lvl + c * cs, nz, e, T);
            }
            atomicAdd(&s_count[j], local);
        }
        __syncthreads();
        if (tid == 0) {
            int first = ncand;  // smallest feasible candidate
            for (int k = ncand - 1; k >= 0; --k)
                if (s_count[k] <= nitems) first = k;
            if (first < ncand) s_hi = lo + step * static_cast<float>(first + 1);
            if (first > 0) s_lo = lo + step * static_cast<float>(first);
            s_first = first < ncand ? first : -1;
        }
        __syncthreads();
        const int f = s_first;
        if (f >= 0) {
            if (ncand > 1) {
                if (tid < ncol) best[tid] = cand[f * ncol + tid];
            } else {
                for (int c = tid; c < ncol; c += blockDim.x)
                    best[c] = greedy_count(lvl + c * cs, nz, e, s_hi);
            }
        }
        __syncthreads();
    }
    const float T = s_hi;

    for (int i = tid; i < kPlanThreads; i += blockDim.x) staged[i] = make_int4(0, 0, 0, 0);
    if (tid == 0) {
        int off = 0;
        for (int c = 0; c < ncol; ++c) {
            const int k = best[c];
            best[c] = off;
            off += k;
        }
        s_total = off;
    }
    __syncthreads();
    if (s_total > nitems) {
        // not reached (every bound kept was tested feasible with the same greedy cut); one
        // chunk per column keeps every cell covered regardless
        for (int c = tid; c < ncol; c += blockDim.x) items[c] = make_int4(c % ntx, c / ntx, 0, nz);
        for (int i = ncol + tid; i < nitems; i += blockDim.x) items[i] = make_int4(0, 0, 0, 0);
        return;
    }
    // chunks are cut column by column into LDS, then ordered by the middle of their z range so
    // that each XCD (a contiguous range of items) takes a z band across all column tiles, as with
    // uniform chunks (measured 2-5% faster than column-major order at 96^3 and 128^3)
    for (int c = tid; c < ncol; c += blockDim.x) {
        const int tix = c % ntx, tiy = c / ntx;
        int o = best[c];
        const int oend = c + 1 < ncol ? best[c + 1] : s_total;
        int z0 = 0;
        float acc = e;
        const uint32_t *col = lvl + c * cs;
        uint32_t word = 0u;
        for (int z = 0; z < nz; ++z) {
            if ((z & 3) == 0) word = col[z >> 2];
            const float w = plane_cost(word, z & 3);
            // the cut count equals best[c] (same greedy, same bound); the guard only keeps the
            // writes inside this column's range, extending its last chunk, should they differ
            if (acc + w > T && o < oend - 1) {
                staged[o++] = make_int4(tix, tiy, z0, z);
                z0 = z;
                acc = e + w;
            } else {
                acc += w;
            }
        }
        staged[o] = make_int4(tix, tiy, z0, nz);
    }
    __syncthreads();
    // key: z0 + z1 (idle items last), then the cut order; bitonic sort over a power of two
    int np2 = 1;
    while (np2 < nitems) np2 <<= 1;
    for (int i = tid; i < np2; i += blockDim.x) {
        uint32_t k = 0xFFFFFFFFu;
        if (i < nitems) k = (static_cast<uint32_t>(i < s_total ? staged[i].z + staged[i].w : 0x3FFFFF) << 10) | i;
        key[i] = k;
    }
    __syncthreads();
    for (int k = 2; k <= np2; k <<= 1) {
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int i = tid; i < np2; i += blockDim.x) {
                const int ixj = i ^ jj;
                if (ixj > i) {
                    const uint32_t a = key[i], b = key[ixj];
                    if ((a > b) == ((i & k) == 0)) {
                        key[i] = b;
                        key[ixj] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < nitems; i += blockDim.x) {
        const int from = static_cast<int>(key[i] & 0x3FFu);
        items[i] = from < s_total ? staged[from] : make_int4(0, 0, 0, 0);
    }
}

}  // namespace

size_t plan_shared(int64_t ncol, int64_t nz) {
    return static_cast<size_t>((16 + 4) * kPlanThreads + 4 * (ncol + kPlanThreads + ncol * (((nz + 3) / 4) | 1)));
}

bool plan_fits(int ntx, int nty, int nz, int nitems) {
    const int64_t ncol = static_cast<int64_t>(ntx) * nty;
    return ncol >= 1 && ncol <= nitems && nitems <= kPlanThreads && nz < (1 << 21) &&
           static_cast<int64_t>(plan_shared(ncol, nz)) <= kPlanMaxShared;
}

hipError_t launch_plan(const uint8_t *wmap, const uint32_t *itime, int ntx, int nty, int nz, int nw, float alpha,
                       float e, int4 *items, int nitems, hipStream_t s) {
    if (!plan_fits(ntx, nty, nz, nitems) || !(e >= 0.0f) || !(alpha >= 0.0f)) return hipErrorInvalidValue;
    const size_t shm = plan_shared(static_cast<int64_t>(ntx) * nty, nz);
    switch (nw) {
    case 8: plan_kernel<8><<<1, kPlanThreads, shm, s>>>(wmap, itime, ntx, nty, nz, alpha, e, items, nitems); break;
    case 4: plan_kernel<4><<<1, kPlanThreads, shm, s>>>(wmap, itime, ntx, nty, nz, alpha, e, items, nitems); break;
    case 2: plan_kernel<2><<<1, kPlanThreads, shm, s>>>(wmap, itime, ntx, nty, nz, alpha, e, items, nitems); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cglbm
