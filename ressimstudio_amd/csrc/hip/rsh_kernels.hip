// HIP kernels of the linear-solve path for gfx950 (MI355X): BSR SpMV over row slices, level-scheduled block
// ILU(0) factorisation and sweeps, the CPR pressure kernels, scalar AMG level kernels, deterministic reductions
// and the fused Krylov vector updates.  Every kernel is HBM- or latency-bound (a 3x3 FP64 block carries 18 flops
// per 76 bytes), so the design goal is coalesced streaming: one block row per lane of a 64-wide wavefront with
// the block entries stored element-major across the 64 rows of a slice (one 512-byte load per entry).
//
// Kernels that restate a reference loop keep its order of floating-point operations and disable contraction
// (#pragma clang fp contract(off)), so that on rows handled by one lane the results equal the reference's
// (bsr.cpp:14-56 block products, :67-87 inversion, :183-215 row elimination, :437-596 sweeps; amg.cpp:98-131).
#include <algorithm>

#include "rsh_internal.hpp"

namespace rsh {

namespace {

constexpr int kBlock = 256;            // 4 wavefronts per workgroup

__device__ inline double waveSum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// a slice / long row i of a batched system that stopped iterating (owner -1: the slice mixes systems, always run)
__device__ inline bool gatedOff(const int* __restrict__ owner, int i, const KScalars* __restrict__ ks) {
    if (!ks || !owner) return false;
    const int o = owner[i];
    return o >= 0 && ks[o].state != kRun;
}
// sum over aligned groups of G lanes (G a power of two <= 64)
template <int G>
__device__ inline double groupSum(double v) {
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, G);
    return v;
}

template <typename T>
__device__ inline T toT(double v) { return static_cast<T>(v); }

// ---------------------------------------------------------------- dense block helpers (registers)
template <int NB>
__device__ inline bool invertBlockDev(double* a) {
#pragma clang fp contract(off)
    double m[NB][2 * NB];
    double scale = 0;
#pragma unroll
    for (int i = 0; i < NB * NB; ++i) scale = fmax(scale, fabs(a[i]));
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) { m[i][j] = a[i * NB + j]; m[i][NB + j] = i == j ? 1.0 : 0.0; }
    if (!(scale > 0) || !isfinite(scale)) return false;
    for (int k = 0; k < NB; ++k) {
        int piv = k;
        for (int i = k + 1; i < NB; ++i) if (fabs(m[i][k]) > fabs(m[piv][k])) piv = i;
        if (fabs(m[piv][k]) <= 1e-14 * scale) return false;
        if (piv != k)
            for (int j = 0; j < 2 * NB; ++j) { const double t = m[k][j]; m[k][j] = m[piv][j]; m[piv][j] = t; }
        const double inv = 1.0 / m[k][k];
        for (int j = 0; j < 2 * NB; ++j) m[k][j] *= inv;
        for (int i = 0; i < NB; ++i) {
            if (i == k || m[i][k] == 0.0) continue;
            const double f = m[i][k];
            for (int j = 0; j < 2 * NB; ++j) m[i][j] -= f * m[k][j];
        }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) a[i * NB + j] = m[i][NB + j];
    return true;
}

// invert a pivot block; singular -> diagonal shift by 1e-8 * scale, else identity / scale (bsr.cpp:202-212)
template <int NB>
__device__ inline int invertPivotDev(double* d) {
#pragma clang fp contract(off)
    double save[NB * NB];
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) save[e] = d[e];
    if (invertBlockDev<NB>(d)) return 0;
    double sc = 0;
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) sc = fmax(sc, fabs(save[e]));
    if (!(sc > 0) || !isfinite(sc)) sc = 1.0;
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) d[e] = save[e];
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        const double x = d[e * NB + e];
        d[e * NB + e] = x + ((x >= 0 ? 1e-8 : -1e-8) * sc + (x == 0 ? sc * 1e-8 : 0));
    }
    if (!invertBlockDev<NB>(d)) {
#pragma unroll
        for (int e = 0; e < NB * NB; ++e) d[e] = 0;
#pragma unroll
        for (int e = 0; e < NB; ++e) d[e * NB + e] = 1.0 / sc;
    }
    return 1;
}

// ---------------------------------------------------------------- layout packing
template <typename TA>
__global__ void kPackSlices(int nb2, const double* __restrict__ src, const int* __restrict__ srcIdx, long long nslots, TA* __restrict__ dst) {
    const long long total = nslots * nb2 * kWave;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total; idx += (long long)gridDim.x * blockDim.x) {
        const int lane = int(idx & 63);
        const long long q = idx >> 6;
        const int e = int(q % nb2);
        const long long slot = q / nb2;
        const int s = srcIdx[slot * kWave + lane];
        dst[idx] = s >= 0 ? toT<TA>(src[(long long)s * nb2 + e]) : TA(0);
    }
}
template <typename TA>
__global__ void kPackLong(int nb2, const double* __restrict__ src, const int* __restrict__ srcIdx, int count, TA* __restrict__ dst) {
    const long long total = (long long)count * nb2;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total; idx += (long long)gridDim.x * blockDim.x) {
        const long long k = idx / nb2;
        const int e = int(idx % nb2);
        const int s = srcIdx[k];
        dst[idx] = s >= 0 ? toT<TA>(src[(long long)s * nb2 + e]) : TA(0);
    }
}
__global__ void kPackColumn0(int nb, const double* __restrict__ src, const int* __restrict__ srcIdx, long long nslots, double* __restrict__ dst) {
    const long long total = nslots * nb * kWave;
    const int nb2 = nb * nb;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total; idx += (long long)gridDim.x * blockDim.x) {
        const int lane = int(idx & 63);
        const long long q = idx >> 6;
        const int e = int(q % nb);
        const long long slot = q / nb;
        const int s = srcIdx[slot * kWave + lane];
        dst[idx] = s >= 0 ? src[(long long)s * nb2 + e * nb] : 0.0;
    }
}
__global__ void kGatherBlocks(int nb2, const double* __restrict__ src, const int* __restrict__ map, int count, double* __restrict__ dst) {
    const long long total = (long long)count * nb2;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total; idx += (long long)gridDim.x * blockDim.x)
        dst[idx] = src[(long long)map[idx / nb2] * nb2 + idx % nb2];
}

inline int gridFor(long long work, int per = kBlock) {
    long long g = (work + per - 1) / per;
    if (g < 1) g = 1;
    if (g > 65535LL * 8) g = 65535LL * 8;
    return int(g);
}

// a slot's column: absolute (CT = int) or relative to the lane's row r (CT = short, the 16-bit layout: half the index
// bytes); NT: non-temporal load (streamed once)
template <typename CT, bool NT>
__device__ inline int slotCol(const CT* __restrict__ col, long long at, int r) {
    const CT c = NT ? __builtin_nontemporal_load(col + at) : col[at];
    if constexpr (sizeof(CT) == sizeof(int)) return int(c);
    else return r + int(c);
}

// ---------------------------------------------------------------- BSR SpMV over slices (+ long rows)
// y_i = sum_p A_p x_col(p) (bsr.cpp:109-126); per lane the blocks are visited in column order and each block
// product is formed as s_e = sum_f a_ef x_f before y_e += s_e, as gemvAdd does.  LN > 1 (laneSliceLayout): LN lanes
// per row, each summing every LN-th block of the row in that way, then the LN partial sums added pairwise (lanes
// xor 1, then xor 2), which every lane of the row ends with.
template <int NB, typename TA, typename CT, int LN = 1>
__global__ __launch_bounds__(kBlock) void kSpmvSliced(int nslices, const int2* __restrict__ info, const int* __restrict__ row,
                                                      const CT* __restrict__ col, const TA* __restrict__ val, int nlong,
                                                      const int* __restrict__ lrow, const int* __restrict__ lptr,
                                                      const int* __restrict__ lcol, const TA* __restrict__ lval,
                                                      const double* __restrict__ x, double* __restrict__ y, const int* __restrict__ owner,
                                                      const int* __restrict__ lowner, const KScalars* __restrict__ gate) {
#pragma clang fp contract(off)
    const int wid = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wid < nslices) {
        if (gatedOff(owner, wid, gate)) return;
        const int2 inf = info[wid];
        const int r = row[(long long)wid * kWave + lane];
        double acc[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[e] = 0;
        // two slots per round: both slots' columns and blocks, then both x gathers, are in flight before the first
        // product; the products are formed and added slot by slot in column order as before
        int k = 0;
        for (; k + 2 <= inf.x; k += 2) {
            const long long slot = (long long)inf.y + k;
            double av[2][NB * NB], xv[2][NB];
            int c[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                // matrix values and columns are streamed once: non-temporal loads keep the caches for the x gathers
                c[u] = slotCol<CT, true>(col, (slot + u) * kWave + lane, r);
                const TA* a = val + (slot + u) * (NB * NB * kWave) + lane;
#pragma unroll
                for (int ef = 0; ef < NB * NB; ++ef) av[u][ef] = double(__builtin_nontemporal_load(a + ef * kWave));
            }
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int f = 0; f < NB; ++f) xv[u][f] = x[(long long)c[u] * NB + f];
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int e = 0; e < NB; ++e) {
                    double s = 0;
#pragma unroll
                    for (int f = 0; f < NB; ++f) s += av[u][e * NB + f] * xv[u][f];
                    acc[e] += s;
                }
        }
        for (; k < inf.x; ++k) {
            const long long slot = (long long)inf.y + k;
            const int c = slotCol<CT, true>(col, slot * kWave + lane, r);
            const TA* a = val + slot * (NB * NB * kWave) + lane;
            double xv[NB];
#pragma unroll
            for (int f = 0; f < NB; ++f) xv[f] = x[(long long)c * NB + f];
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                double s = 0;
#pragma unroll
                for (int f = 0; f < NB; ++f) s += double(__builtin_nontemporal_load(a + (e * NB + f) * kWave)) * xv[f];
                acc[e] += s;
            }
        }
        if constexpr (LN > 1) {
#pragma unroll
            for (int m = 1; m < LN; m <<= 1)
#pragma unroll
                for (int e = 0; e < NB; ++e) acc[e] += __shfl_xor(acc[e], m, 64);
        }
        if (r >= 0 && (lane & (LN - 1)) == 0) {
#pragma unroll
            for (int e = 0; e < NB; ++e) y[(long long)r * NB + e] = acc[e];
        }
        return;
    }
    const int q = wid - nslices;
    if (q >= nlong || gatedOff(lowner, q, gate)) return;
    double part[NB];
#pragma unroll
    for (int e = 0; e < NB; ++e) part[e] = 0;
    for (int p = lptr[q] + lane; p < lptr[q + 1]; p += kWave) {
        const int c = lcol[p];
        const TA* a = lval + (long long)p * NB * NB;
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            double s = 0;
#pragma unroll
            for (int f = 0; f < NB; ++f) s += double(a[e * NB + f]) * x[(long long)c * NB + f];
            part[e] += s;
        }
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) part[e] = waveSum(part[e]);
    if (lane == 0) {
        const int r = lrow[q];
#pragma unroll
        for (int e = 0; e < NB; ++e) y[(long long)r * NB + e] = part[e];
    }
}

// ---------------------------------------------------------------- BILU0 level sweeps
// One lane's slots of a sweep: acc_e -= sum_f a_ef y_col,f per block in slot (column) order, s_e formed first as
// in the block products of bsr.cpp:14-56.  kSweepSlots slots per round: their columns and blocks, then their
// gathers of y, are in flight before the first product; the arithmetic is the same slot by slot.
constexpr int kSweepSlots = 3;
template <int NB, typename TA, typename TV>
__device__ inline void sliceRowSweep(const int* __restrict__ col, const TA* __restrict__ val, int2 inf, int lane, const TV* y, TV (&acc)[NB]) {
#pragma clang fp contract(off)
    constexpr int U = kSweepSlots;
    int k = 0;
    for (; k + U <= inf.x; k += U) {
        const long long slot = (long long)inf.y + k;
        TV av[U][NB * NB], yk[U][NB];
        int c[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = __builtin_nontemporal_load(col + (slot + u) * kWave + lane);
            const TA* a = val + (slot + u) * (NB * NB * kWave) + lane;
#pragma unroll
            for (int ef = 0; ef < NB * NB; ++ef) av[u][ef] = TV(__builtin_nontemporal_load(a + ef * kWave));
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int f = 0; f < NB; ++f) yk[u][f] = y[(long long)c[u] * NB + f];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                TV s = 0;
#pragma unroll
                for (int f = 0; f < NB; ++f) s += av[u][e * NB + f] * yk[u][f];
                acc[e] -= s;
            }
    }
    for (; k < inf.x; ++k) {
        const long long slot = (long long)inf.y + k;
        const int c = __builtin_nontemporal_load(col + slot * kWave + lane);
        const TA* a = val + slot * (NB * NB * kWave) + lane;
        TV yk[NB];
#pragma unroll
        for (int f = 0; f < NB; ++f) yk[f] = y[(long long)c * NB + f];
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            TV s = 0;
#pragma unroll
            for (int f = 0; f < NB; ++f) s += TV(__builtin_nontemporal_load(a + (e * NB + f) * kWave)) * yk[f];
            acc[e] -= s;
        }
    }
}

// Forward (bsr.cpp:575-583): y_i = r_i - sum_{k<i} L_ik y_k, y in the working (level-ordered) numbering, r read
// through the original row index.  TA: factor storage, TV: working-vector type (accumulation type).
template <int NB, typename TA, typename TV>
__global__ __launch_bounds__(kBlock) void kIluForward(int s0, int ns, int q0, int nq, const int2* __restrict__ info,
                                                      const int* __restrict__ row, const int* __restrict__ orig,
                                                      const int* __restrict__ col, const TA* __restrict__ val,
                                                      const int* __restrict__ lrow, const int* __restrict__ lorig,
                                                      const int* __restrict__ lptr, const int* __restrict__ lcol,
                                                      const TA* __restrict__ lval, const double* __restrict__ r, TV* __restrict__ y,
                                                      const int* __restrict__ owner, const int* __restrict__ lowner,
                                                      const KScalars* __restrict__ gate, const int* __restrict__ fdsrc,
                                                      const TA* __restrict__ fdval, double* __restrict__ z, const double* __restrict__ xp) {
#pragma clang fp contract(off)
    const int wid = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wid < ns) {
        const int sl = s0 + wid;
        if (gatedOff(owner, sl, gate)) return;
        const int wr = row[(long long)sl * kWave + lane];
        if (wr < 0) return;
        const int o = orig[(long long)sl * kWave + lane];
        const int2 inf = info[sl];
        TV acc[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[e] = TV(r[(long long)o * NB + e]);
        sliceRowSweep<NB, TA, TV>(col, val, inf, lane, y, acc);
        if (fdval && fdsrc[(long long)sl * kWave + lane] >= 0) {
            // no upper entries and no later reader: the backward step z = D^-1 y now (kIluBackward's arithmetic)
            const TA* d = fdval + (long long)sl * (NB * NB * kWave) + lane;
            TV t[NB];
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                TV sacc = 0;
#pragma unroll
                for (int f = 0; f < NB; ++f) sacc += TV(d[(e * NB + f) * kWave]) * acc[f];
                t[e] = sacc;
            }
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                y[(long long)wr * NB + e] = t[e];
                double v = double(t[e]);
                if (e == 0 && xp) v += xp[o];
                z[(long long)o * NB + e] = v;
            }
            return;
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) y[(long long)wr * NB + e] = acc[e];
        return;
    }
    const int qi = wid - ns;
    if (qi >= nq) return;
    const int q = q0 + qi;
    if (gatedOff(lowner, q, gate)) return;
    TV part[NB];
#pragma unroll
    for (int e = 0; e < NB; ++e) part[e] = 0;
    for (int p = lptr[q] + lane; p < lptr[q + 1]; p += kWave) {
        const int c = lcol[p];
        const TA* a = lval + (long long)p * NB * NB;
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            TV s = 0;
#pragma unroll
            for (int f = 0; f < NB; ++f) s += TV(a[e * NB + f]) * y[(long long)c * NB + f];
            part[e] += s;
        }
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) part[e] = TV(waveSum(double(part[e])));
    if (lane == 0) {
        const int wr = lrow[q], o = lorig[q];
#pragma unroll
        for (int e = 0; e < NB; ++e) y[(long long)wr * NB + e] = TV(r[(long long)o * NB + e]) - part[e];
    }
}

// Backward (bsr.cpp:584-595): z_i = U_ii^-1 (y_i - sum_{j>i} U_ij z_j), in place in y; the result is also
// scattered to z in the original numbering, with xp added to the pressure unknown when given (the CPR
// correction z_i0 += x_p,i, bsr.cpp:1083).
template <int NB, typename TA, typename TV>
__global__ __launch_bounds__(kBlock) void kIluBackward(int s0, int ns, int q0, int nq, const int2* __restrict__ info,
                                                       const int* __restrict__ row, const int* __restrict__ orig,
                                                       const int* __restrict__ col, const TA* __restrict__ val,
                                                       const TA* __restrict__ dval, const int* __restrict__ lrow,
                                                       const int* __restrict__ lorig, const int* __restrict__ lptr,
                                                       const int* __restrict__ lcol, const TA* __restrict__ lval,
                                                       const TA* __restrict__ ldval, TV* __restrict__ y, double* __restrict__ z,
                                                       const double* __restrict__ xp, const int* __restrict__ owner,
                                                       const int* __restrict__ lowner, const KScalars* __restrict__ gate, int writeY) {
#pragma clang fp contract(off)
    const int wid = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    TV acc[NB];
    int wr, o;
    const TA* d;
    int dstride;
    if (wid < ns) {
        const int sl = s0 + wid;
        if (gatedOff(owner, sl, gate)) return;
        wr = row[(long long)sl * kWave + lane];
        if (wr < 0) return;
        o = orig[(long long)sl * kWave + lane];
        const int2 inf = info[sl];
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[e] = y[(long long)wr * NB + e];
        sliceRowSweep<NB, TA, TV>(col, val, inf, lane, y, acc);
        d = dval + (long long)sl * (NB * NB * kWave) + lane;
        dstride = kWave;
    } else {
        const int qi = wid - ns;
        if (qi >= nq) return;
        const int q = q0 + qi;
        if (gatedOff(lowner, q, gate)) return;
        TV part[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e) part[e] = 0;
        for (int p = lptr[q] + lane; p < lptr[q + 1]; p += kWave) {
            const int c = lcol[p];
            const TA* a = lval + (long long)p * NB * NB;
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                TV s = 0;
#pragma unroll
                for (int f = 0; f < NB; ++f) s += TV(a[e * NB + f]) * y[(long long)c * NB + f];
                part[e] += s;
            }
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) part[e] = TV(waveSum(double(part[e])));
        if (lane != 0) return;
        wr = lrow[q];
        o = lorig[q];
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[e] = y[(long long)wr * NB + e] - part[e];
        d = ldval + (long long)q * NB * NB;
        dstride = 1;
    }
    TV t[NB];
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        TV s = 0;
#pragma unroll
        for (int f = 0; f < NB; ++f) s += TV(d[(e * NB + f) * dstride]) * acc[f];
        t[e] = s;
    }
    if (writeY) {   // the last launch of the sweep has no later reader of y
#pragma unroll
        for (int e = 0; e < NB; ++e) y[(long long)wr * NB + e] = t[e];
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        double v = double(t[e]);
        if (e == 0 && xp) v += xp[o];
        z[(long long)o * NB + e] = v;
    }
}


// Red / black tile sweeps: one wavefront per tile of one colour (kTileWaveTiles tiles per workgroup) walks the tile's
// slices - one per local level, natural order inside the tile - forward in order or backward in reverse, with a
// workgroup barrier between levels, which makes a level's writes to the working vector visible to the next one.  Each
// row's arithmetic is kIluForward's (y_i = r_i - sum L_ik y_k, slot by slot) or kIluBackward's (z_i = U_ii^-1 (y_i -
// sum U_ik z_k)), with the values of the rows of other tiles (the other colour, the rows behind the tiles) read from
// the working vector as those kernels do.  tiles[t] = (first slice, end slice, first working row, end row).
// (Measured: holding the tile's own values in LDS and fetching every slot of a level in one round, with the next
// level's columns prefetched, needed 246-256 registers, two wavefronts per SIMD, and was slower.)
constexpr int kTileWaveTiles = 4;
template <int NB, typename TA, typename TV, bool FWD>
__global__ __launch_bounds__(kTileWaveTiles * kWave) void kIluTileWave(int ntiles, const int4* __restrict__ tiles, const int2* __restrict__ info,
                                                                       const int* __restrict__ row, const int* __restrict__ orig,
                                                                       const int* __restrict__ col, const TA* __restrict__ val,
                                                                       const TA* __restrict__ dval, const double* __restrict__ r, TV* y,
                                                                       double* __restrict__ z, const double* __restrict__ xp) {
#pragma clang fp contract(off)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tb = blockIdx.x * kTileWaveTiles;
    int steps = 0;
    for (int k = 0; k < kTileWaveTiles; ++k)
        if (tb + k < ntiles) {
            const int4 Tk = tiles[tb + k];
            steps = max(steps, Tk.y - Tk.x);
        }
    int4 T = make_int4(0, 0, 0, 0);
    if (tb + wave < ntiles) T = tiles[tb + wave];
    const int ns = T.y - T.x;
    for (int q = 0; q < steps; ++q) {
        if (q < ns) {
            const int sl = FWD ? T.x + q : T.y - 1 - q;
            const int wr = row[(long long)sl * kWave + lane];
            if (wr >= 0) {
                const int o = orig[(long long)sl * kWave + lane];
                const int2 inf = info[sl];
                TV acc[NB];
#pragma unroll
                for (int e = 0; e < NB; ++e) acc[e] = FWD ? TV(r[(long long)o * NB + e]) : y[(long long)wr * NB + e];
                sliceRowSweep<NB, TA, TV>(col, val, inf, lane, y, acc);
                if (FWD) {
#pragma unroll
                    for (int e = 0; e < NB; ++e) y[(long long)wr * NB + e] = acc[e];
                } else {
                    const TA* d = dval + (long long)sl * (NB * NB * kWave) + lane;
                    TV t[NB];
#pragma unroll
                    for (int e = 0; e < NB; ++e) {
                        TV sacc = 0;
#pragma unroll
                        for (int f = 0; f < NB; ++f) sacc += TV(d[(e * NB + f) * kWave]) * acc[f];
                        t[e] = sacc;
                    }
#pragma unroll
                    for (int e = 0; e < NB; ++e) {
                        y[(long long)wr * NB + e] = t[e];
                        double v = double(t[e]);
                        if (e == 0 && xp) v += xp[o];
                        z[(long long)o * NB + e] = v;
                    }
                }
            }
        }
        __syncthreads();
    }
}


// ---------------------------------------------------------------- BILU0 factorisation, one level, in place
// Row i (bsr.cpp:183-215): for each k < i in row order: L_ik = A_ik U_kk^-1; A_ij -= L_ik U_kj for every j > k
// present in row i; then the pivot block is inverted (U_ii^-1 stored).  Rows of one level only read rows of
// earlier levels.
template <int NB>
__global__ void kBilu0Level(const int* __restrict__ rows, int r0, int r1, const int* __restrict__ rp, const int* __restrict__ ci,
                            const int* __restrict__ diag, double* __restrict__ val, int* __restrict__ nreg, int stop, int lean) {
#pragma clang fp contract(off)
    const int idx = r0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= r1) return;
    const int i = rows ? rows[idx] : idx;
    const int pe = rp[i + 1], pd = diag[i];
    constexpr int B2 = NB * NB;
    for (int p = rp[i]; p < pd; ++p) {
        const int k = ci[p];
        if (stop >= 0 && k >= stop) break;   // the rest of the row is eliminated by kIluGroupFactor
        double* ap = val + (long long)p * B2;
        const double* dk = val + (long long)diag[k] * B2;
        double L[B2];
#pragma unroll
        for (int ii = 0; ii < NB; ++ii) {
            double rr[NB];
#pragma unroll
            for (int j = 0; j < NB; ++j) rr[j] = 0;
#pragma unroll
            for (int kk = 0; kk < NB; ++kk) {
                const double f = ap[ii * NB + kk];
#pragma unroll
                for (int j = 0; j < NB; ++j) rr[j] += f * dk[kk * NB + j];
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) L[ii * NB + j] = rr[j];
        }
#pragma unroll
        for (int e = 0; e < B2; ++e) ap[e] = L[e];
        int s = p + 1;
        const int qe = rp[k + 1];
        for (int q = diag[k] + 1; q < qe && s < pe; ++q) {
            const int j = ci[q];
            while (s < pe && ci[s] < j) ++s;
            if (s < pe && ci[s] == j) {
                const double* u = val + (long long)q * B2;
                double* as = val + (long long)s * B2;
#pragma unroll
                for (int ii = 0; ii < NB; ++ii) {
                    double rr[NB];
#pragma unroll
                    for (int jj = 0; jj < NB; ++jj) rr[jj] = 0;
#pragma unroll
                    for (int kk = 0; kk < NB; ++kk) {
                        const double f = L[ii * NB + kk];
#pragma unroll
                        for (int jj = 0; jj < NB; ++jj) rr[jj] += f * u[kk * NB + jj];
                    }
#pragma unroll
                    for (int jj = 0; jj < NB; ++jj) as[ii * NB + jj] -= rr[jj];
                }
            }
        }
    }
    if (stop >= 0) return;
    double d[B2];
    double* dp = val + (long long)pd * B2;
#pragma unroll
    for (int e = 0; e < B2; ++e) d[e] = dp[e];
    const int reg = invertPivotDev<NB>(d);
#pragma unroll
    for (int e = 0; e < B2; ++e) dp[e] = d[e];
    if (reg && nreg) atomicAdd(nreg, 1);
    // lean (C45, bsr.cpp:216-263): the finished row is stored rounded to float32; later rows read it rounded
    if (lean)
        for (long long e = (long long)rp[i] * B2; e < (long long)pe * B2; ++e) val[e] = double(float(val[e]));
}
// lean factorisation: the matrix rounded to float32 before the elimination
__global__ void kRoundToFloat(long long n, double* __restrict__ x) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        x[i] = double(float(x[i]));
}

// The rows of a group coupled inside a run of small levels (a well's perforated cells), once the contributions
// of the rows before the run are in (kBilu0Level with stop), are eliminated right-looking: for every member k
// in ascending order, invert its pivot, form L_ik = A_ik U_kk^-1 for the later members i, then A_ij -= L_ik U_kj
// for the later members i, j.  Every entry receives the same updates in the same order and with the same block
// arithmetic as the row-by-row elimination of bilu0Row (bsr.cpp:183-215), so the factors are bit-identical; the
// dependent steps are one per member instead of one merge walk per entry pair.  One workgroup of kGroupThreads
// per group (a step's trailing update is (K - t - 1)^2 independent block pairs); idx[a * K + b] is the factor
// block of members (a, b) or -1.
constexpr int kGroupThreads = 1024;
template <int NB>
__global__ __launch_bounds__(kGroupThreads) void kIluGroupFactor(const int* __restrict__ gstart, const int* __restrict__ grows,
                                                       const long long* __restrict__ gidxOff, const int* __restrict__ gidx,
                                                       const int* __restrict__ diag, double* val, int* __restrict__ nreg,
                                                       const int* __restrict__ rp, int lean) {
#pragma clang fp contract(off)
    constexpr int B2 = NB * NB;
    const int g = blockIdx.x;
    const int m0 = gstart[g], K = gstart[g + 1] - m0;
    const int* idx = gidx + gidxOff[g];
    for (int t = 0; t < K; ++t) {
        const int k = grows[m0 + t];
        double* dk = val + (long long)diag[k] * B2;
        if (threadIdx.x == 0) {
            double d[B2];
#pragma unroll
            for (int e = 0; e < B2; ++e) d[e] = dk[e];
            const int reg = invertPivotDev<NB>(d);
#pragma unroll
            for (int e = 0; e < B2; ++e) dk[e] = d[e];
            if (reg && nreg) atomicAdd(nreg, 1);
        }
        __syncthreads();
        if (lean) {
            // member k's row is finished (its earlier members' updates and its pivot): stored rounded to float32
            // before any later member reads it (C45, bsr.cpp:216-263)
            for (long long e = (long long)rp[k] * B2 + threadIdx.x; e < (long long)rp[k + 1] * B2; e += blockDim.x) val[e] = double(float(val[e]));
            __syncthreads();
        }
        // L_ak = A_ak U_kk^-1 for the later members a
        for (int a = t + 1 + int(threadIdx.x); a < K; a += int(blockDim.x)) {
            const int p = idx[(long long)a * K + t];
            if (p < 0) continue;
            double* ap = val + (long long)p * B2;
            double L[B2];
#pragma unroll
            for (int ii = 0; ii < NB; ++ii) {
                double rr[NB];
#pragma unroll
                for (int j = 0; j < NB; ++j) rr[j] = 0;
#pragma unroll
                for (int kk = 0; kk < NB; ++kk) {
                    const double f = ap[ii * NB + kk];
#pragma unroll
                    for (int j = 0; j < NB; ++j) rr[j] += f * dk[kk * NB + j];
                }
#pragma unroll
                for (int j = 0; j < NB; ++j) L[ii * NB + j] = rr[j];
            }
#pragma unroll
            for (int e = 0; e < B2; ++e) ap[e] = L[e];
        }
        __syncthreads();
        // A_ab -= L_ak U_kb for the later members a, b (entries of the pattern only)
        const int r = K - t - 1;
        for (int pr = int(threadIdx.x); pr < r * r; pr += int(blockDim.x)) {
            const int a = t + 1 + pr / r, b = t + 1 + pr % r;
            const int pl = idx[(long long)a * K + t], pu = idx[(long long)t * K + b], pab = idx[(long long)a * K + b];
            if (pl < 0 || pu < 0 || pab < 0) continue;
            const double* L = val + (long long)pl * B2;
            const double* u = val + (long long)pu * B2;
            double* as = val + (long long)pab * B2;
#pragma unroll
            for (int ii = 0; ii < NB; ++ii) {
                double rr[NB];
#pragma unroll
                for (int jj = 0; jj < NB; ++jj) rr[jj] = 0;
#pragma unroll
                for (int kk = 0; kk < NB; ++kk) {
                    const double f = L[ii * NB + kk];
#pragma unroll
                    for (int jj = 0; jj < NB; ++jj) rr[jj] += f * u[kk * NB + jj];
                }
#pragma unroll
                for (int jj = 0; jj < NB; ++jj) as[ii * NB + jj] -= rr[jj];
            }
        }
        __syncthreads();
    }
}


// ---------------------------------------------------------------- fused runs of small levels
// Levels with few rows (the colour classes of the well cliques, the coarse AMG levels) are swept by one
// workgroup of 1024 threads that walks the run of levels in order with a workgroup barrier between levels:
// one launch replaces one launch per level, and the barrier costs far less than a dispatch.
template <typename TA>
struct SliceView {
    const int2* info; const int* row; const int* orig; const int* col; const TA* val; const TA* dval;
    const int* lrow; const int* lorig; const int* lptr; const int* lcol; const TA* lval; const TA* ldval;
    const int* owner; const int* lowner;
};

constexpr int kFuseThreads = 1024;

// A run of small levels [l0, l1) whose rows are the contiguous working rows [wbase, wbase + wcount), swept by one
// workgroup.  Phase A (all rows at once): the contributions of entries whose column lies outside the run - already
// final before the run starts - and the right-hand side, into LDS.  Phase B (level by level, a workgroup barrier
// between levels): the entries inside the run, read from and written to LDS.  Phase C: the run's values to HBM.
// The sequential part thus touches only the run's own couplings (the well cliques) and LDS.
template <int NB, typename TA, typename TV, bool FWD, bool INSIDE>
__device__ inline void runTask(const SliceView<TA>& A, int code, int lane, int wbase, int wend, TV* ys, const double* __restrict__ r,
                               const TV* y) {
#pragma clang fp contract(off)
    auto use = [&](int c) { return INSIDE ? (c >= wbase && c < wend) : !(c >= wbase && c < wend); };
    auto yval = [&](int c, int f) -> TV { return INSIDE ? ys[(c - wbase) * NB + f] : y[(long long)c * NB + f]; };
    if (code >= 0) {
        const int wr = A.row[(long long)code * kWave + lane];
        if (wr < 0) return;
        const int2 inf = A.info[code];
        TV acc[NB];
        if (INSIDE) {
#pragma unroll
            for (int e = 0; e < NB; ++e) acc[e] = ys[(wr - wbase) * NB + e];
        } else {
            const int o = A.orig[(long long)code * kWave + lane];
#pragma unroll
            for (int e = 0; e < NB; ++e) acc[e] = FWD ? TV(r[(long long)o * NB + e]) : y[(long long)wr * NB + e];
        }
        for (int k = 0; k < inf.x; ++k) {
            const long long slot = (long long)inf.y + k;
            const int c = A.col[slot * kWave + lane];
            if (!use(c)) continue;
            const TA* a = A.val + slot * (NB * NB * kWave) + lane;
            TV yk[NB];
#pragma unroll
            for (int f = 0; f < NB; ++f) yk[f] = yval(c, f);
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                TV sacc = 0;
#pragma unroll
                for (int f = 0; f < NB; ++f) sacc += TV(a[(e * NB + f) * kWave]) * yk[f];
                acc[e] -= sacc;
            }
        }
        if (INSIDE && !FWD) {
            const TA* d = A.dval + (long long)code * (NB * NB * kWave) + lane;
            TV tv[NB];
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                TV sacc = 0;
#pragma unroll
                for (int f = 0; f < NB; ++f) sacc += TV(d[(e * NB + f) * kWave]) * acc[f];
                tv[e] = sacc;
            }
#pragma unroll
            for (int e = 0; e < NB; ++e) acc[e] = tv[e];
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) ys[(wr - wbase) * NB + e] = acc[e];
        return;
    }
    const int q = -code - 1;
    TV part[NB];
#pragma unroll
    for (int e = 0; e < NB; ++e) part[e] = 0;
    for (int p = A.lptr[q] + lane; p < A.lptr[q + 1]; p += kWave) {
        const int c = A.lcol[p];
        if (!use(c)) continue;
        const TA* a = A.lval + (long long)p * NB * NB;
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            TV sacc = 0;
#pragma unroll
            for (int f = 0; f < NB; ++f) sacc += TV(a[e * NB + f]) * yval(c, f);
            part[e] += sacc;
        }
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) part[e] = TV(waveSum(double(part[e])));
    if (lane != 0) return;
    const int wr = A.lrow[q];
    TV acc[NB];
    if (INSIDE) {
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[e] = ys[(wr - wbase) * NB + e] - part[e];
        if (!FWD) {
            const TA* d = A.ldval + (long long)q * NB * NB;
            TV tv[NB];
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                TV sacc = 0;
#pragma unroll
                for (int f = 0; f < NB; ++f) sacc += TV(d[e * NB + f]) * acc[f];
                tv[e] = sacc;
            }
#pragma unroll
            for (int e = 0; e < NB; ++e) acc[e] = tv[e];
        }
    } else {
        const int o = A.lorig[q];
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[e] = (FWD ? TV(r[(long long)o * NB + e]) : y[(long long)wr * NB + e]) - part[e];
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) ys[(wr - wbase) * NB + e] = acc[e];
}

template <int NB, typename TA, typename TV, bool FWD>
__global__ __launch_bounds__(kFuseThreads) void kIluFused(int l0, int l1, const int* __restrict__ sStart, const int* __restrict__ qStart,
                                                          SliceView<TA> A, int wbase, int wcount, const int* __restrict__ perm,
                                                          const double* __restrict__ r, TV* y, double* __restrict__ z,
                                                          const double* __restrict__ xp) {
    __shared__ TV ys[kRunRowsMax * NB];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    constexpr int waves = kFuseThreads / kWave;
    const int wend = wbase + wcount;
    // phase A: every task of the run, contributions from outside the run
    {
        const int t0 = sStart[l0], t1 = sStart[l1], u0 = qStart[l0], u1 = qStart[l1];
        const int ns = t1 - t0, nq = u1 - u0;
        for (int t = wave; t < ns + nq; t += waves)
            runTask<NB, TA, TV, FWD, false>(A, t < ns ? t0 + t : -(u0 + t - ns) - 1, lane, wbase, wend, ys, r, y);
    }
    __syncthreads();
    // phase B: level by level, the couplings inside the run
    for (int li = 0; li < l1 - l0; ++li) {
        const int l = FWD ? l0 + li : l1 - 1 - li;
        const int s0 = sStart[l], ns = sStart[l + 1] - s0, q0 = qStart[l], nq = qStart[l + 1] - q0;
        for (int t = wave; t < ns + nq; t += waves)
            runTask<NB, TA, TV, FWD, true>(A, t < ns ? s0 + t : -(q0 + t - ns) - 1, lane, wbase, wend, ys, r, y);
        __syncthreads();
    }
    // phase C: results to HBM (the working vector, and z in the original numbering after the backward sweep)
    for (int k = threadIdx.x; k < wcount; k += kFuseThreads) {
#pragma unroll
        for (int e = 0; e < NB; ++e) y[(long long)(wbase + k) * NB + e] = ys[k * NB + e];
        if (!FWD) {
            const int o = perm[wbase + k];
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                double v = double(ys[k * NB + e]);
                if (e == 0 && xp) v += xp[o];
                z[(long long)o * NB + e] = v;
            }
        }
    }
}

// ---------------------------------------------------------------- tiled natural-order sweeps
// The natural-order sweeps have one dependent level per grid diagonal (363 at 60 x 220 x 85).  Here the rows are
// grouped into units (a compact tile of the grid at one tile level, rows ordered by local level): a unit depends only
// on units of earlier tile levels and on its own rows, so one launch per tile level runs one workgroup per unit, which
// walks the unit's local levels with workgroup barriers, the unit's values in LDS.
// Bulk rows (one per lane of a slice wavefront): phase A loads the row's first TileCap slots into registers, gathers
// the values of the slots outside the unit (final before the launch) and applies the leading outside slots; phase B, at
// the row's local level, applies the remaining slots in slot order (inside values from LDS).  Every bulk row thus keeps
// the slot order and arithmetic of the level-scheduled sweep.  Long rows (a well's perforated cells after the Schur
// fill): one wavefront takes the unit's long rows one after another, lanes over slots; phase A sums the outside slots
// and stages the inside slots (index and block) in LDS, phase B sums those at the row's local level.
constexpr int kTileThreads = kTileWaves * kWave;
constexpr int kTileLongWaves = 4;   // wavefronts that share a unit's long rows in phase A (when the slices leave them)
template <int NB> struct TileCap { static constexpr int v = NB <= 2 ? 6 : (NB == 3 ? 3 : 2); };

struct TileLds { int blk, ldv, ys, lpart, red, lidx, lv, llev, lnin, lwr, total; };
__host__ __device__ inline int alignLds(int b) { return (b + 15) & ~15; }
// dynamic LDS of a tile launch: staged long-row blocks and inverted pivots, the unit's values, long-row partial sums,
// the long-row wavefront's per-lane partials, staged inside indices, local level starts, long-row levels / inside
// counts / working rows
__host__ __device__ inline TileLds tileLds(int rowsCap, int levelsCap, int longMax, int longCap, int nb, int tvBytes, int taBytes) {
    TileLds L{};
    int o = 0;
    L.blk = o; o = alignLds(o + longMax * longCap * nb * nb * taBytes);
    L.ldv = o; o = alignLds(o + longMax * nb * nb * taBytes);
    L.ys = o; o = alignLds(o + rowsCap * nb * tvBytes);
    L.lpart = o; o = alignLds(o + longMax * nb * tvBytes);
    L.red = o; o = alignLds(o + kTileLongWaves * kWave * nb * tvBytes);
    L.lidx = o; o = alignLds(o + longMax * longCap * 4);
    L.lv = o; o = alignLds(o + (levelsCap + 1) * 4);
    L.llev = o; o = alignLds(o + longMax * 4);
    L.lnin = o; o = alignLds(o + longMax * 4);
    L.lwr = o; o = alignLds(o + longMax * 4);
    L.total = o;
    return L;
}

// sum of red[j * NB + e] over j in [0, cnt) in ascending order (four loads in flight at a time)
template <int NB, typename TV>
__device__ inline void ldsOrderedSum(const TV* red, int cnt, TV (&sum)[NB]) {
#pragma clang fp contract(off)
    for (int j0 = 0; j0 < cnt; j0 += 4) {
        TV t[4][NB];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (j0 + u < cnt) {
#pragma unroll
                for (int e = 0; e < NB; ++e) t[u][e] = red[(j0 + u) * NB + e];
            }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (j0 + u < cnt) {
#pragma unroll
                for (int e = 0; e < NB; ++e) sum[e] += t[u][e];
            }
    }
}

template <int NB, typename TA, typename TV, bool FWD>
__global__ __launch_bounds__(kTileThreads) void kIluTile(int u0, const int* __restrict__ unitStart, const int* __restrict__ sStart,
                                                         const int* __restrict__ qStart, const int* __restrict__ lvPtr,
                                                         const int* __restrict__ lvStart, SliceView<TA> A, int rowsCap, int levelsCap,
                                                         int longMax, int longCap, const int* __restrict__ perm,
                                                         const double* __restrict__ r, TV* y, double* __restrict__ z,
                                                         const double* __restrict__ xp) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smem[];
    const TileLds Ls = tileLds(rowsCap, levelsCap, longMax, longCap, NB, int(sizeof(TV)), int(sizeof(TA)));
    TA* lblk = reinterpret_cast<TA*>(smem + Ls.blk);
    TA* ldv = reinterpret_cast<TA*>(smem + Ls.ldv);
    TV* ys = reinterpret_cast<TV*>(smem + Ls.ys);
    TV* lpart = reinterpret_cast<TV*>(smem + Ls.lpart);
    TV* red = reinterpret_cast<TV*>(smem + Ls.red);
    int* lidx = reinterpret_cast<int*>(smem + Ls.lidx);
    int* lv = reinterpret_cast<int*>(smem + Ls.lv);
    int* llev = reinterpret_cast<int*>(smem + Ls.llev);
    int* lnin = reinterpret_cast<int*>(smem + Ls.lnin);
    int* lwr = reinterpret_cast<int*>(smem + Ls.lwr);
    constexpr int CAP = TileCap<NB>::v;
    constexpr int NB2 = NB * NB;
    const int u = u0 + blockIdx.x;
    const int wb = unitStart[u], we = unitStart[u + 1];
    const int s0 = sStart[u], ns = sStart[u + 1] - s0, q0 = qStart[u], nq = qStart[u + 1] - q0;
    const int lp = lvPtr[u], nl = lvPtr[u + 1] - lp - 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // the slice's row and shape are loaded together with the level starts (one round trip before the barrier)
    int wr = -1, o = -1;
    int2 inf = make_int2(0, 0);
    if (wave < ns) {
        wr = A.row[(long long)(s0 + wave) * kWave + lane];
        o = A.orig[(long long)(s0 + wave) * kWave + lane];
        inf = A.info[s0 + wave];
    }
    for (int t = threadIdx.x; t <= nl; t += kTileThreads) lv[t] = lvStart[lp + t];
    __syncthreads();
    // local level of the unit's row with local index li (li < the unit's row count = lv[nl])
    auto levelOf = [&](int li) {
        int l = 0;
        while (lv[l + 1] <= li) ++l;
        return l;
    };
    auto inside = [&](int c) { return c >= wb && c < we; };

    // ---- phase A
    int myLev = -1, width = 0, kfirst = 0;
    long long slot0 = 0;
    TV acc[NB];
    int idx[CAP];       // LDS row of an inside slot, -1 otherwise
    TA av[CAP][NB2];
    TV yv[CAP][NB];     // outside slots: their block products (formed here); inside slots: the values read at the step
    TA dv[NB2];
#pragma unroll
    for (int k = 0; k < CAP; ++k) {
        idx[k] = -1;
#pragma unroll
        for (int ef = 0; ef < NB2; ++ef) av[k][ef] = 0;
#pragma unroll
        for (int e = 0; e < NB; ++e) yv[k][e] = 0;
    }
    if (wave < ns) {
        const int sl = s0 + wave;
        if (wr >= 0) {
            width = inf.x;
            slot0 = inf.y;
            myLev = levelOf(wr - wb);
            if (FWD) {
#pragma unroll
                for (int e = 0; e < NB; ++e) acc[e] = TV(r[(long long)o * NB + e]);
            } else {
#pragma unroll
                for (int e = 0; e < NB; ++e) acc[e] = y[(long long)wr * NB + e];
            }
            int c[CAP];
#pragma unroll
            for (int k = 0; k < CAP; ++k)
                if (k < width) {
                    c[k] = __builtin_nontemporal_load(A.col + (slot0 + k) * kWave + lane);
                    idx[k] = inside(c[k]) ? c[k] - wb : -1;
                    const TA* a = A.val + (slot0 + k) * (NB2 * kWave) + lane;
#pragma unroll
                    for (int ef = 0; ef < NB2; ++ef) av[k][ef] = __builtin_nontemporal_load(a + ef * kWave);
                }
            if (!FWD) {
                const TA* d = A.dval + (long long)sl * (NB2 * kWave) + lane;
#pragma unroll
                for (int ef = 0; ef < NB2; ++ef) dv[ef] = __builtin_nontemporal_load(d + ef * kWave);
            }
            TV yk[CAP][NB];
#pragma unroll
            for (int k = 0; k < CAP; ++k)
                if (k < width && idx[k] < 0) {
#pragma unroll
                    for (int f = 0; f < NB; ++f) yk[k][f] = y[(long long)c[k] * NB + f];
                }
            // the block products of the outside slots (s_e = sum_f a_ef y_f, f ascending); the leading ones applied
#pragma unroll
            for (int k = 0; k < CAP; ++k)
                if (k < width && idx[k] < 0) {
#pragma unroll
                    for (int e = 0; e < NB; ++e) yv[k][e] = 0;
#pragma unroll
                    for (int f = 0; f < NB; ++f)
#pragma unroll
                        for (int e = 0; e < NB; ++e) yv[k][e] += TV(av[k][e * NB + f]) * yk[k][f];
                }
#pragma unroll
            for (int k = 0; k < CAP; ++k)
                if (k < width && kfirst == k && idx[k] < 0) {
#pragma unroll
                    for (int e = 0; e < NB; ++e) acc[e] -= yv[k][e];
                    kfirst = k + 1;
                }
            if (kfirst == CAP) {
                for (int k = CAP; k < width; ++k) {
                    const int ck = A.col[(slot0 + k) * kWave + lane];
                    if (inside(ck)) break;
                    const TA* a = A.val + (slot0 + k) * (NB2 * kWave) + lane;
#pragma unroll
                    for (int e = 0; e < NB; ++e) {
                        TV sacc = 0;
#pragma unroll
                        for (int f = 0; f < NB; ++f) sacc += TV(a[(e * NB + f) * kWave]) * y[(long long)ck * NB + f];
                        acc[e] -= sacc;
                    }
                    kfirst = k + 1;
                }
            }
        }
    } else if (wave >= ns && wave < ns + kTileLongWaves && nq > 0) {
        // the unit's long rows, spread over the wavefronts left after the slices (up to kTileLongWaves), G lanes each:
        // outside slots summed (per-lane partials added in lane order by the group's first lane), inside slots staged
        const int lw = wave - ns;
        const int nlw = min(kTileLongWaves, kTileWaves - ns);
        const int mine = lw < nq ? (nq - lw + nlw - 1) / nlw : 0;   // long rows lw, lw + nlw, ... of this wavefront
        int G = kWave;
        while (G > 1 && G * mine > kWave) G >>= 1;
        const int grp = lane / G, gl = lane % G;
        const int qi = grp < mine ? lw + grp * nlw : -1;
        TV* myred = red + lw * kWave * NB;
        TV part[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e) part[e] = 0;
        int p0 = 0, p1 = 0;
        if (qi >= 0) {
            p0 = A.lptr[q0 + qi];
            p1 = A.lptr[q0 + qi + 1];
        }
        const unsigned long long gmask = (G == kWave ? ~0ull : ((1ull << G) - 1ull)) << (grp * G);
        int base = 0;
        int trips = 0;
        for (int j = 0; j < mine; ++j) {
            const int qq = q0 + lw + j * nlw;
            trips = max(trips, (A.lptr[qq + 1] - A.lptr[qq] + G - 1) / G);
        }
        constexpr int UT = 4;   // trips whose loads are in flight together
        for (int t0 = 0; t0 < trips; t0 += UT) {
            int cc[UT];
            bool valid[UT], in[UT];
            TA ab[UT][NB2];
            TV yk[UT][NB];
#pragma unroll
            for (int v = 0; v < UT; ++v) {
                const int p = p0 + (t0 + v) * G + gl;
                valid[v] = qi >= 0 && t0 + v < trips && p < p1;
                cc[v] = valid[v] ? A.lcol[p] : -1;
            }
#pragma unroll
            for (int v = 0; v < UT; ++v) {
                in[v] = valid[v] && inside(cc[v]);
                if (valid[v]) {
                    const TA* a = A.lval + (long long)(p0 + (t0 + v) * G + gl) * NB2;
#pragma unroll
                    for (int ef = 0; ef < NB2; ++ef) ab[v][ef] = a[ef];
                }
                if (valid[v] && !in[v]) {
#pragma unroll
                    for (int f = 0; f < NB; ++f) yk[v][f] = y[(long long)cc[v] * NB + f];
                }
            }
#pragma unroll
            for (int v = 0; v < UT; ++v) {
                if (valid[v] && !in[v]) {
#pragma unroll
                    for (int e = 0; e < NB; ++e) {
                        TV sacc = 0;
#pragma unroll
                        for (int f = 0; f < NB; ++f) sacc += TV(ab[v][e * NB + f]) * yk[v][f];
                        part[e] += sacc;
                    }
                }
                const unsigned long long m = __ballot(in[v] ? 1 : 0) & gmask;
                if (in[v]) {
                    const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
                    lidx[qi * longCap + pos] = cc[v] - wb;
#pragma unroll
                    for (int ef = 0; ef < NB2; ++ef) lblk[(qi * longCap + pos) * NB2 + ef] = ab[v][ef];
                }
                base += __popcll(m);
            }
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) myred[lane * NB + e] = part[e];
        if (!FWD && qi >= 0)
            for (int t = gl; t < NB2; t += G) ldv[qi * NB2 + t] = A.ldval[(long long)(q0 + qi) * NB2 + t];
        __builtin_amdgcn_wave_barrier();
        if (qi >= 0 && gl == 0) {
            TV sum[NB];
#pragma unroll
            for (int e = 0; e < NB; ++e) sum[e] = 0;
            ldsOrderedSum<NB, TV>(myred + grp * G * NB, G, sum);
            const int q = q0 + qi;
            const int w = A.lrow[q];
            lwr[qi] = w;
            llev[qi] = levelOf(w - wb);
            lnin[qi] = base;
            if (FWD) {
                const int o = A.lorig[q];
#pragma unroll
                for (int e = 0; e < NB; ++e) lpart[qi * NB + e] = TV(r[(long long)o * NB + e]) - sum[e];
            } else {
#pragma unroll
                for (int e = 0; e < NB; ++e) lpart[qi * NB + e] = y[(long long)w * NB + e] - sum[e];
            }
        }
    }
    __syncthreads();

    // ---- phase B: local level by level
    // the long rows are in local-level order: the wavefront keeps the next one to take (FWD: ascending)
    int qcur = FWD ? 0 : nq - 1;
    int nextLev = -1;
    if (wave == ns && nq > 0) nextLev = llev[qcur];
    for (int li = 0; li < nl; ++li) {
        const int l = FWD ? li : nl - 1 - li;
        if (wr >= 0 && myLev == l) {
            // the inside values first (their reads in flight together), their block products (independent chains side
            // by side, each summed f ascending), then every remaining slot's product subtracted in slot order
            // (straight-line: a slot that takes no part reads LDS row 0 and subtracts an exact zero, x - 0 = x)
            TV yk[CAP][NB], sk[CAP][NB];
#pragma unroll
            for (int k = 0; k < CAP; ++k) {
                const int ii = idx[k] >= 0 ? idx[k] : 0;
#pragma unroll
                for (int f = 0; f < NB; ++f) yk[k][f] = ys[ii * NB + f];
            }
#pragma unroll
            for (int k = 0; k < CAP; ++k)
#pragma unroll
                for (int e = 0; e < NB; ++e) sk[k][e] = 0;
#pragma unroll
            for (int f = 0; f < NB; ++f)
#pragma unroll
                for (int k = 0; k < CAP; ++k)
#pragma unroll
                    for (int e = 0; e < NB; ++e) sk[k][e] += TV(av[k][e * NB + f]) * yk[k][f];
#pragma unroll
            for (int k = 0; k < CAP; ++k) {
                const bool take = k >= kfirst && k < width;
#pragma unroll
                for (int e = 0; e < NB; ++e) acc[e] -= take ? (idx[k] >= 0 ? sk[k][e] : yv[k][e]) : TV(0);
            }
            for (int k = kfirst > CAP ? kfirst : CAP; k < width; ++k) {
                const int ck = A.col[(slot0 + k) * kWave + lane];
                const TA* a = A.val + (slot0 + k) * (NB2 * kWave) + lane;
                TV yk[NB];
                if (inside(ck)) {
#pragma unroll
                    for (int f = 0; f < NB; ++f) yk[f] = ys[(ck - wb) * NB + f];
                } else {
#pragma unroll
                    for (int f = 0; f < NB; ++f) yk[f] = y[(long long)ck * NB + f];
                }
#pragma unroll
                for (int e = 0; e < NB; ++e) {
                    TV sacc = 0;
#pragma unroll
                    for (int f = 0; f < NB; ++f) sacc += TV(a[(e * NB + f) * kWave]) * yk[f];
                    acc[e] -= sacc;
                }
            }
            if (!FWD) {
                TV tv[NB];
#pragma unroll
                for (int e = 0; e < NB; ++e) tv[e] = 0;
#pragma unroll
                for (int f = 0; f < NB; ++f)
#pragma unroll
                    for (int e = 0; e < NB; ++e) tv[e] += TV(dv[e * NB + f]) * acc[f];
#pragma unroll
                for (int e = 0; e < NB; ++e) acc[e] = tv[e];
            }
#pragma unroll
            for (int e = 0; e < NB; ++e) ys[(wr - wb) * NB + e] = acc[e];
        }
        while (nextLev == l) {
            // long row qcur: lanes over its staged inside slots, per-lane partials summed in lane order by lane 0
            const int qi = qcur;
            const int cnt = lnin[qi];
            TV part[NB];
#pragma unroll
            for (int e = 0; e < NB; ++e) part[e] = 0;
            for (int j = lane; j < cnt; j += kWave) {
                const int li2 = lidx[qi * longCap + j];
                const TA* a = lblk + (qi * longCap + j) * NB2;
#pragma unroll
                for (int e = 0; e < NB; ++e) {
                    TV sacc = 0;
#pragma unroll
                    for (int f = 0; f < NB; ++f) sacc += TV(a[e * NB + f]) * ys[li2 * NB + f];
                    part[e] += sacc;
                }
            }
#pragma unroll
            for (int e = 0; e < NB; ++e) red[lane * NB + e] = part[e];
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) {
                TV sum[NB];
#pragma unroll
                for (int e = 0; e < NB; ++e) sum[e] = 0;
                ldsOrderedSum<NB, TV>(red, cnt < kWave ? cnt : kWave, sum);
                TV a2[NB];
#pragma unroll
                for (int e = 0; e < NB; ++e) a2[e] = lpart[qi * NB + e] - sum[e];
                if (!FWD) {
                    TV tv[NB];
#pragma unroll
                    for (int e = 0; e < NB; ++e) {
                        TV sacc = 0;
#pragma unroll
                        for (int f = 0; f < NB; ++f) sacc += TV(ldv[qi * NB2 + e * NB + f]) * a2[f];
                        tv[e] = sacc;
                    }
#pragma unroll
                    for (int e = 0; e < NB; ++e) a2[e] = tv[e];
                }
#pragma unroll
                for (int e = 0; e < NB; ++e) ys[(lwr[qi] - wb) * NB + e] = a2[e];
            }
            __builtin_amdgcn_wave_barrier();
            qcur += FWD ? 1 : -1;
            nextLev = (qcur >= 0 && qcur < nq) ? llev[qcur] : -1;
        }
        __syncthreads();
    }

    // ---- phase C: the unit's values to HBM (and z in the original numbering after the backward sweep)
    const int cnt = we - wb;
    for (int k = threadIdx.x; k < cnt; k += kTileThreads) {
#pragma unroll
        for (int e = 0; e < NB; ++e) y[(long long)(wb + k) * NB + e] = ys[k * NB + e];
        if (!FWD) {
            const int o = perm[wb + k];
#pragma unroll
            for (int e = 0; e < NB; ++e) {
                double v = double(ys[k * NB + e]);
                if (e == 0 && xp) v += xp[o];
                z[(long long)o * NB + e] = v;
            }
        }
    }
}

// ---------------------------------------------------------------- dense groups inside a run of small levels
// Rows of a run that are coupled to each other inside the run form groups (a well's perforated cells, coupled
// all-to-all by the Schur-reduced bottom-hole-pressure row, make one group per well).  Per group the unit
// block-lower factor L_GG (forward sweep) and (I + D^-1 U_GG)^-1 D^-1 (backward sweep, D^-1 the inverted pivots)
// are inverted once per factorisation into dense column-major matrices, so that a sweep over the run is one
// parallel pass (the contributions from rows outside the run) and one dense product per group, instead of one
// dependent step per level.

// outside contributions of every task of the levels [l0, l1): rt[(row - wbase) * NB + e], one wavefront per task
template <int NB, typename TA, typename TV, bool FWD>
__global__ __launch_bounds__(256) void kRunOutside(int t0, int ns, int u0, int nq, SliceView<TA> A, int wbase, int wend,
                                                   const double* __restrict__ r, const TV* y, TV* rt, const KScalars* __restrict__ gate) {
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (t >= ns + nq) return;
    if (t < ns ? gatedOff(A.owner, t0 + t, gate) : gatedOff(A.lowner, u0 + t - ns, gate)) return;
    runTask<NB, TA, TV, FWD, false>(A, t < ns ? t0 + t : -(u0 + t - ns) - 1, lane, wbase, wend, rt, r, y);
}

// group g: rows rows[start[g] .. start[g+1]) (working numbering, ascending), dense K x K at off[g], K = k NB, column-major
// with exact zeros outside the block triangle.  One workgroup per 64-row tile of a group: its 16 wavefronts split the
// tile's non-zero columns, each lane one row; the partial sums are added in a fixed order.  With ycopy given (forward
// only) the result is also written in the run's local numbering, where the backward product of the same run reads
// its input when the run's rows have no upper couplings outside the run (the backward outside pass would only copy y).
template <int NB, typename TM, typename TV, bool FWD>
__global__ __launch_bounds__(kDenseThreads) void kRunDense(const int2* __restrict__ tasks, const int* __restrict__ start,
                                                          const int* __restrict__ rows, const long long* __restrict__ off,
                                                          const TM* __restrict__ inv, int wbase, const TV* __restrict__ rt, TV* y,
                                                          double* __restrict__ z, const int* __restrict__ perm,
                                                          const double* __restrict__ xp, const int* __restrict__ gowner,
                                                          const KScalars* __restrict__ gate, TV* __restrict__ ycopy) {
    constexpr int W = kDenseThreads / kWave;
    __shared__ TV v[kDenseMaxRows * NB];
    __shared__ TV part[W][kWave];
    const int2 tk = tasks[blockIdx.x];   // x group, y first scalar row of the tile
    if (gatedOff(gowner, tk.x, gate)) return;
    const int r0 = start[tk.x], K = (start[tk.x + 1] - r0) * NB;
    for (int t = threadIdx.x; t < K; t += kDenseThreads) v[t] = rt[(long long)(rows[r0 + t / NB] - wbase) * NB + t % NB];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int a = tk.y + lane;
    const int aLast = min(K, tk.y + kWave) - 1;
    const int c0 = FWD ? 0 : (tk.y / NB) * NB, c1 = FWD ? (aLast / NB + 1) * NB : K;
    TV s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    if (a < K) {
        const TM* M = inv + off[tk.x] + a;
        int b = c0 + wave;
        for (; b + 3 * W < c1; b += 4 * W) {
            s0 += TV(M[(long long)b * K]) * v[b];
            s1 += TV(M[(long long)(b + W) * K]) * v[b + W];
            s2 += TV(M[(long long)(b + 2 * W) * K]) * v[b + 2 * W];
            s3 += TV(M[(long long)(b + 3 * W) * K]) * v[b + 3 * W];
        }
        for (; b < c1; b += W) s0 += TV(M[(long long)b * K]) * v[b];
    }
    part[wave][lane] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (wave != 0 || a >= K) return;
    TV out = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) out += part[w][lane];
    const int blk = a / NB, w = rows[r0 + blk], e = a - blk * NB;
    y[(long long)w * NB + e] = out;
    if (FWD && ycopy) ycopy[(w - wbase) * NB + e] = out;
    if (!FWD) {
        const int o = perm[w];
        double val = double(out);
        if (e == 0 && xp) val += xp[o];
        z[(long long)o * NB + e] = val;
    }
}

// gather: raw[group, local row i, local col j] = lu block src (map = {group, i, j, src}); raw zeroed beforehand
__global__ void kDenseGather(int count, int nb, const int4* __restrict__ map, const int* __restrict__ start,
                             const long long* __restrict__ off, const double* __restrict__ lu, double* __restrict__ raw) {
    const int nb2 = nb * nb;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)count * nb2) return;
    const int q = int(t / nb2), ef = int(t % nb2), e = ef / nb, f = ef % nb;
    const int4 m = map[q];
    const long long K = (long long)(start[m.x + 1] - start[m.x]) * nb;
    raw[off[m.x] + (m.z * nb + f) * K + m.y * nb + e] = lu[(long long)m.w * nb2 + ef];
}

// dense NB x NB blocks (i, j) of a K x K column-major matrix
template <int NB>
__device__ inline void denseBlk(const double* B, long long K, int i, int j, double (&o)[NB][NB]) {
#pragma unroll
    for (int e = 0; e < NB; ++e)
#pragma unroll
        for (int f = 0; f < NB; ++f) o[e][f] = B[(j * NB + f) * K + i * NB + e];
}
template <int NB>
__device__ inline void densePut(double* B, long long K, int i, int j, const double (&o)[NB][NB]) {
#pragma unroll
    for (int e = 0; e < NB; ++e)
#pragma unroll
        for (int f = 0; f < NB; ++f) B[(j * NB + f) * K + i * NB + e] = o[e][f];
}
template <int NB>
__device__ inline void denseMul(const double (&a)[NB][NB], const double (&b)[NB][NB], double (&o)[NB][NB]) {
#pragma unroll
    for (int e = 0; e < NB; ++e)
#pragma unroll
        for (int f = 0; f < NB; ++f) {
            double s = 0;
#pragma unroll
            for (int h = 0; h < NB; ++h) s += a[e][h] * b[h][f];
            o[e][f] = s;
        }
}
// upper sweep, before its inversion: W_im = D^-1_i U_im (i < m) in place; grid (blocks over k^2, groups)
template <int NB>
__global__ __launch_bounds__(256) void kDenseScaleU(const int* __restrict__ start, const long long* __restrict__ off, double* raw) {
    const int g = blockIdx.y;
    const int k = start[g + 1] - start[g];
    const long long K = (long long)k * NB;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= k * k) return;
    const int i = t % k, m = t / k;
    if (i >= m) return;
    double* M = raw + off[g];
    double d[NB][NB], u[NB][NB], w[NB][NB];
    denseBlk<NB>(M, K, i, i, d);
    denseBlk<NB>(M, K, i, m, u);
    denseMul<NB>(d, u, w);
    densePut<NB>(M, K, i, m, w);
}
// X = L^-1 (lower) or X = (I + W)^-1 D^-1 (upper, W from kDenseScaleU; D^-1 on raw's diagonal blocks).  Grid
// (P, groups): workgroup p of a group owns the block columns j = p, p + P, ...; the elimination never couples two
// columns of X (X_ij -= M_im X_mj), so each block gets the same updates in the same order as in one workgroup.
template <int NB, bool UPPER>
__global__ __launch_bounds__(256) void kDenseInvert(const int* __restrict__ start, const long long* __restrict__ off, const double* raw,
                                                    double* inv) {
    const int g = blockIdx.y, p = blockIdx.x, P = gridDim.x;
    const int k = start[g + 1] - start[g];
    if (p >= k) return;                                  // the whole workgroup
    const int nc = (k - p + P - 1) / P;                  // block columns owned
    const long long K = (long long)k * NB;
    const double* M = raw + off[g];
    double* X = inv + off[g];
    for (long long t = threadIdx.x; t < (long long)nc * NB * K; t += blockDim.x) {
        const long long c = t / (NB * K), rem = t % (NB * K);
        const long long col = (p + c * P) * NB + rem / K, row = rem % K;
        X[col * K + row] = row == col ? 1.0 : 0.0;
    }
    __syncthreads();
    // right-looking elimination: step m finalises block row m of X, then removes its coupling from the rows after
    // it (lower: rows i > m, columns j <= m; upper: rows i < m, columns j >= m)
    for (int s = 0; s < k; ++s) {
        const int m = UPPER ? k - 1 - s : s;
        int c0, na, ni, i0;
        if (UPPER) {
            c0 = m <= p ? 0 : (m - p + P - 1) / P;
            na = nc - c0;
            ni = m;
            i0 = 0;
        } else {
            c0 = 0;
            na = m >= p ? (m - p) / P + 1 : 0;
            ni = k - 1 - m;
            i0 = m + 1;
        }
        if (na <= 0 || ni <= 0) continue;                // uniform over the workgroup
        for (int t = threadIdx.x; t < ni * na; t += blockDim.x) {
            const int i = i0 + t / na;
            const int j = p + (c0 + t % na) * P;
            double a[NB][NB], x[NB][NB], q[NB][NB], c[NB][NB];
            denseBlk<NB>(M, K, i, m, a);
            denseBlk<NB>(X, K, m, j, x);
            denseMul<NB>(a, x, q);
            denseBlk<NB>(X, K, i, j, c);
#pragma unroll
            for (int e = 0; e < NB; ++e)
#pragma unroll
                for (int f = 0; f < NB; ++f) c[e][f] -= q[e][f];
            densePut<NB>(X, K, i, j, c);
        }
        __syncthreads();
    }
    if (UPPER) {
        // X_ij <- X_ij D^-1_j (i <= j)
        for (int t = threadIdx.x; t < k * nc; t += blockDim.x) {
            const int i = t % k, j = p + (t / k) * P;
            if (i > j) continue;
            double x[NB][NB], d[NB][NB], o[NB][NB];
            denseBlk<NB>(X, K, i, j, x);
            denseBlk<NB>(M, K, j, j, d);
            denseMul<NB>(x, d, o);
            densePut<NB>(X, K, i, j, o);
        }
    }
}

// synthetic Newton iterations of a served model: every block row i scaled by (1 + amp * u_i), u_i uniform in
// [-1, 1) from a SplitMix64 hash of (seed, i) (new accumulation / mobility values per cell; the pattern and the
// character of the Jacobian stay those of the model)
__global__ void kPerturbRows(int n, const int* __restrict__ rp, int nb2, const double* __restrict__ base, double* __restrict__ out, double amp,
                             unsigned long long seed) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long z = seed * 0x9E3779B97F4A7C15ull + (unsigned long long)i + 1ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = double(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    const double f = 1.0 + amp * u;
    for (long long k = (long long)rp[i] * nb2; k < (long long)rp[i + 1] * nb2; ++k) out[k] = base[k] * f;
}

// ---------------------------------------------------------------- block-Jacobi second stage (DeviceCprAmg, bsr.cpp:1304-1311, 1359-1367)
// D_b^-1 per block row; a singular block falls back to the inverted absolute diagonal, as the reference does
template <int NB>
__global__ void kBlockDiagInv(int n, const int* __restrict__ diag, const double* __restrict__ val, double* __restrict__ out) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* D = val + (long long)diag[i] * NB * NB;
    double a[NB * NB];
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) a[e] = D[e];
    if (!invertBlockDev<NB>(a)) {
#pragma unroll
        for (int e = 0; e < NB * NB; ++e) a[e] = 0;
#pragma unroll
        for (int e = 0; e < NB; ++e) a[e * NB + e] = 1.0 / fmax(1e-30, fabs(D[e * NB + e]));
    }
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) out[(long long)i * NB * NB + e] = a[e];
}
// z_i = D_i^-1 t_i, then z_i0 += xp_i (z1 = E xp); TD: storage of D^-1
template <int NB, typename TD>
__global__ void kBlockJacobi(int n, const TD* __restrict__ dinv, const double* __restrict__ t, const double* __restrict__ xp, double* __restrict__ z) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const TD* D = dinv + (long long)i * NB * NB;
    double ti[NB];
#pragma unroll
    for (int f = 0; f < NB; ++f) ti[f] = t[(long long)i * NB + f];
#pragma unroll
    for (int e = 0; e < NB; ++e) {
        double s = 0;
#pragma unroll
        for (int f = 0; f < NB; ++f) s += double(D[e * NB + f]) * ti[f];
        z[(long long)i * NB + e] = (e == 0 ? xp[i] : 0.0) + s;
    }
}

// ---------------------------------------------------------------- well Schur reduction on the device (WellSchur, bsr.cpp:733-990)
// The scalar Jacobian J: cells in blocks of nv unknowns (rows / columns < nc = na nv), then one BHP row per well.
// dinv[w] = 1 / J_ww (refused: |J_ww| <= 1e-12 max |J_w,cells| or non-finite -> bad = 1)
// 1 / d_w per well; bad[0]: a well without a usable BHP derivative, bad[1]: a nonzero well-well entry
__global__ void kWellDinv(int nw, const int* __restrict__ wdiag, const int* __restrict__ wcPtr, const int* __restrict__ wcSrc,
                          const int* __restrict__ wwPtr, const int* __restrict__ wwSrc, const double* __restrict__ jval,
                          double* __restrict__ dinv, int* __restrict__ bad) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    for (int q = wwPtr[w]; q < wwPtr[w + 1]; ++q)
        if (jval[wwSrc[q]] != 0.0) bad[1] = 1;
    const double d = wdiag[w] >= 0 ? jval[wdiag[w]] : 0.0;
    double rowMax = 0;
    for (int q = wcPtr[w]; q < wcPtr[w + 1]; ++q) rowMax = fmax(rowMax, fabs(jval[wcSrc[q]]));
    if (!(fabs(d) > 1e-12 * fmax(rowMax, 1e-300)) || !isfinite(d)) { bad[0] = 1; dinv[w] = 0.0; return; }
    dinv[w] = 1.0 / d;
}
// S value t = (J cell entry mapped to t, if any) - its rank-1 terms (J_cw dinv_w) J_wc in ascending well order, the
// engine's per-entry order (bsr.cpp:810-825: the cell entry, then the wells of the scalar row in column order)
__global__ void kSchurValues(long long nvals, const int* __restrict__ direct, const int* __restrict__ termPtr, const int4* __restrict__ terms,
                             const double* __restrict__ jval, const double* __restrict__ dinv, double* __restrict__ sval) {
#pragma clang fp contract(off)
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nvals) return;
    double acc = 0.0;
    const int d = direct[t];
    if (d >= 0) acc += jval[d];
    for (int k = termPtr[t]; k < termPtr[t + 1]; ++k) {
        const int4 q = terms[k];                 // x: J index of the cell -> well entry, y: well, z: J index of well -> cell
        const double f = jval[q.x] * dinv[q.y];
        acc -= f * jval[q.z];
    }
    sval[t] = acc;
}
// bc_r = b_r - sum_w (J_rw dinv_w) b_w (reduceRhs, bsr.cpp:973-980)
__global__ void kSchurRhs(int nc, const int* __restrict__ cwPtr, const int* __restrict__ cwWell, const int* __restrict__ cwSrc,
                          const double* __restrict__ jval, const double* __restrict__ dinv, const double* __restrict__ b, double* __restrict__ bc) {
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nc) return;
    double s = b[r];
    for (int p = cwPtr[r]; p < cwPtr[r + 1]; ++p) s -= jval[cwSrc[p]] * dinv[cwWell[p]] * b[nc + cwWell[p]];
    bc[r] = s;
}
// x = (x_c, x_w), x_w = (b_w - sum J_wc x_c) dinv_w (recover, bsr.cpp:982-990)
__global__ void kSchurRecover(int nc, int nw, const int* __restrict__ wcPtr, const int* __restrict__ wcCol, const int* __restrict__ wcSrc,
                              const double* __restrict__ jval, const double* __restrict__ dinv, const double* __restrict__ b,
                              const double* __restrict__ xc, double* __restrict__ x) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) x[i] = xc[i];
    if (i >= nw) return;
    double s = b[nc + i];
    for (int p = wcPtr[i]; p < wcPtr[i + 1]; ++p) s -= jval[wcSrc[p]] * xc[wcCol[p]];
    x[nc + i] = s * dinv[i];
}

__global__ void kToFloat(long long n, const double* __restrict__ a, float* __restrict__ b) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = float(a[i]);
}

// Gauss-Seidel over a run of small levels (rows listed level by level, starts in lstart)
__global__ __launch_bounds__(kFuseThreads) void kGsFused(int l0, int l1, const int* __restrict__ lstart, const int* __restrict__ rows,
                                                         const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ diag,
                                                         const double* __restrict__ v, const double* __restrict__ b,
                                                         const double* __restrict__ xin, double* xout, int forward) {
#pragma clang fp contract(off)
    for (int l = l0; l < l1; ++l) {
        for (int t = lstart[l] + threadIdx.x; t < lstart[l + 1]; t += kFuseThreads) {
            const int i = rows[t];
            double s = b[i];
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                const int j = ci[p];
                if (j == i) continue;
                const bool useOut = forward ? (j < i) : (j > i);
                if (useOut) s -= v[p] * xout[j];
                else if (xin) s -= v[p] * xin[j];
            }
            xout[i] = s / v[diag[i]];
        }
        __syncthreads();
    }
}

// BILU0 elimination of long rows (the perforated cells after the Schur fill): one wavefront per row; the L entries
// are taken in order, and for each of them the lanes share the U entries of row k (each finds its column in row
// i by binary search and updates a distinct block), so the updates are the same as in the one-thread loop.
template <int NB>
__global__ __launch_bounds__(kWave) void kBilu0LongRows(const int* __restrict__ rows, int r0, const int* __restrict__ rp,
                                                        const int* __restrict__ ci, const int* __restrict__ diag, double* val,
                                                        int* __restrict__ nreg, int lean) {
#pragma clang fp contract(off)
    constexpr int B2 = NB * NB;
    const int i = rows[r0 + blockIdx.x];
    const int lane = threadIdx.x;
    const int pe = rp[i + 1], pd = diag[i];
    for (int p = rp[i]; p < pd; ++p) {
        const int k = ci[p];
        const double* dk = val + (long long)diag[k] * B2;
        double L[B2];
        {
            const double* ap = val + (long long)p * B2;
#pragma unroll
            for (int ii = 0; ii < NB; ++ii) {
                double rr[NB];
#pragma unroll
                for (int j = 0; j < NB; ++j) rr[j] = 0;
#pragma unroll
                for (int kk = 0; kk < NB; ++kk) {
                    const double f = ap[ii * NB + kk];
#pragma unroll
                    for (int j = 0; j < NB; ++j) rr[j] += f * dk[kk * NB + j];
                }
#pragma unroll
                for (int j = 0; j < NB; ++j) L[ii * NB + j] = rr[j];
            }
        }
        __syncthreads();
        if (lane == 0) {
#pragma unroll
            for (int e = 0; e < B2; ++e) val[(long long)p * B2 + e] = L[e];
        }
        const int qe = rp[k + 1];
        for (int q = diag[k] + 1 + lane; q < qe; q += kWave) {
            const int j = ci[q];
            int lo = p + 1, hi = pe - 1, s = -1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const int cm = ci[mid];
                if (cm == j) { s = mid; break; }
                if (cm < j) lo = mid + 1; else hi = mid - 1;
            }
            if (s < 0) continue;
            const double* u = val + (long long)q * B2;
            double* as = val + (long long)s * B2;
#pragma unroll
            for (int ii = 0; ii < NB; ++ii) {
                double rr[NB];
#pragma unroll
                for (int jj = 0; jj < NB; ++jj) rr[jj] = 0;
#pragma unroll
                for (int kk = 0; kk < NB; ++kk) {
                    const double f = L[ii * NB + kk];
#pragma unroll
                    for (int jj = 0; jj < NB; ++jj) rr[jj] += f * u[kk * NB + jj];
                }
#pragma unroll
                for (int jj = 0; jj < NB; ++jj) as[ii * NB + jj] -= rr[jj];
            }
        }
        __syncthreads();
    }
    if (lane == 0) {
        double d[B2];
        double* dp = val + (long long)pd * B2;
#pragma unroll
        for (int e = 0; e < B2; ++e) d[e] = dp[e];
        const int reg = invertPivotDev<NB>(d);
#pragma unroll
        for (int e = 0; e < B2; ++e) dp[e] = d[e];
        if (reg && nreg) atomicAdd(nreg, 1);
    }
    if (lean) {
        // the finished row stored rounded to float32 (lean, C45)
        __syncthreads();
        for (long long e = (long long)rp[i] * B2 + lane; e < (long long)pe * B2; e += kWave) val[e] = double(float(val[e]));
    }
}

// The same elimination through a precomputed update map, the row held in LDS (natural order, where the well rows are
// long).  For long row t (rows[t] = i) and its lower entry p, the pairs map[mptr[mrow[t] + p - rp[i]] ..] = {q, s}
// list the U blocks q of row k = ci[p] that meet the row, s being the entry they update (offset from rp[i]).  The rows
// before i are final, so the next step's pivot and U blocks are loaded while the current step runs; the sequential
// chain is the row itself, in LDS.  Every entry receives the same updates in the same order with the same block
// arithmetic as kBilu0LongRows: the factors are bit-identical.
template <int NB>
__global__ __launch_bounds__(kWave) void kBilu0LongMapped(const int* __restrict__ rows, int r0, const int* __restrict__ rp,
                                                          const int* __restrict__ ci, const int* __restrict__ diag, double* val,
                                                          int* __restrict__ nreg, int lean, const int* __restrict__ mrow,
                                                          const int* __restrict__ mptr, const int2* __restrict__ map) {
#pragma clang fp contract(off)
    extern __shared__ __align__(16) unsigned char smemRow[];
    double* rowv = reinterpret_cast<double*>(smemRow);
    constexpr int B2 = NB * NB;
    const int t = r0 + blockIdx.x;
    const int i = rows[t];
    const int lane = threadIdx.x;
    const int pb = rp[i], pe = rp[i + 1], pd = diag[i];
    const int m0 = mrow[t];
    for (int e = lane; e < (pe - pb) * B2; e += kWave) rowv[e] = val[(long long)pb * B2 + e];
    // one step's operands: the pivot inverse of row k = ci[p] and this lane's first pair
    auto fetch = [&](int p, double (&dk)[B2], double (&u)[B2], int& sl, int& cnt) {
        const int k = ci[p];
        const double* d = val + (long long)diag[k] * B2;
#pragma unroll
        for (int e = 0; e < B2; ++e) dk[e] = d[e];
        const int a = mptr[m0 + (p - pb)];
        cnt = mptr[m0 + (p - pb) + 1] - a;
        sl = -1;
        if (lane < cnt) {
            const int2 qs = map[a + lane];
            sl = qs.y;
            const double* uu = val + (long long)qs.x * B2;
#pragma unroll
            for (int e = 0; e < B2; ++e) u[e] = uu[e];
        }
    };
    double dkC[B2], uC[B2];
    int sC = -1, cntC = 0;
#pragma unroll
    for (int e = 0; e < B2; ++e) { dkC[e] = 0; uC[e] = 0; }
    if (pb < pd) fetch(pb, dkC, uC, sC, cntC);
    __syncthreads();
    for (int p = pb; p < pd; ++p) {
        double dkN[B2], uN[B2];
        int sN = -1, cntN = 0;
#pragma unroll
        for (int e = 0; e < B2; ++e) { dkN[e] = 0; uN[e] = 0; }
        if (p + 1 < pd) fetch(p + 1, dkN, uN, sN, cntN);
        // L_ik = A_ik U_kk^-1
        double L[B2];
        {
            const double* ap = rowv + (p - pb) * B2;
#pragma unroll
            for (int ii = 0; ii < NB; ++ii) {
                double rr[NB];
#pragma unroll
                for (int j = 0; j < NB; ++j) rr[j] = 0;
#pragma unroll
                for (int kk = 0; kk < NB; ++kk) {
                    const double f = ap[ii * NB + kk];
#pragma unroll
                    for (int j = 0; j < NB; ++j) rr[j] += f * dkC[kk * NB + j];
                }
#pragma unroll
                for (int j = 0; j < NB; ++j) L[ii * NB + j] = rr[j];
            }
        }
        __syncthreads();
        if (lane == 0) {
#pragma unroll
            for (int e = 0; e < B2; ++e) rowv[(p - pb) * B2 + e] = L[e];
        }
        // A_is -= L_ik U_kj for the pairs of this step (distinct s > p)
        auto update = [&](int sl, const double (&u)[B2]) {
            double* as = rowv + sl * B2;
#pragma unroll
            for (int ii = 0; ii < NB; ++ii) {
                double rr[NB];
#pragma unroll
                for (int jj = 0; jj < NB; ++jj) rr[jj] = 0;
#pragma unroll
                for (int kk = 0; kk < NB; ++kk) {
                    const double f = L[ii * NB + kk];
#pragma unroll
                    for (int jj = 0; jj < NB; ++jj) rr[jj] += f * u[kk * NB + jj];
                }
#pragma unroll
                for (int jj = 0; jj < NB; ++jj) as[ii * NB + jj] -= rr[jj];
            }
        };
        if (sC >= 0) update(sC, uC);
        if (cntC > kWave) {
            const int a = mptr[m0 + (p - pb)];
            for (int j = lane + kWave; j < cntC; j += kWave) {
                const int2 qs = map[a + j];
                double u[B2];
#pragma unroll
                for (int e = 0; e < B2; ++e) u[e] = val[(long long)qs.x * B2 + e];
                update(qs.y, u);
            }
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < B2; ++e) { dkC[e] = dkN[e]; uC[e] = uN[e]; }
        sC = sN;
        cntC = cntN;
    }
    if (lane == 0) {
        double d[B2];
        double* dp = rowv + (pd - pb) * B2;
#pragma unroll
        for (int e = 0; e < B2; ++e) d[e] = dp[e];
        const int reg = invertPivotDev<NB>(d);
#pragma unroll
        for (int e = 0; e < B2; ++e) dp[e] = d[e];
        if (reg && nreg) atomicAdd(nreg, 1);
    }
    __syncthreads();
    // the finished row to HBM (lean, C45: stored rounded to float32)
    for (int e = lane; e < (pe - pb) * B2; e += kWave) {
        const double v = rowv[e];
        val[(long long)pb * B2 + e] = lean ? double(float(v)) : v;
    }
}

// ---------------------------------------------------------------- CPR (bsr.cpp:993-1084)
// quasi-IMPES weights (bsr.cpp:1000-1030): w_i = first row of D_ii^-1; singular -> 1 / sum_e D[e][0]; non-finite or
// zero -> 1.  With acc (n accumulation blocks B_i), true-IMPES: B_i is inverted and, if its 1-norm condition estimate
// |B|_1 |B^-1|_1 is at most 1e10 (or nb = 1), the first row of B_i^-1 is the weight; a singular scalar block is used
// as it is; otherwise the cell falls back to the quasi-IMPES weight and is counted in fallbacks.
template <int NB>
__global__ void kCprWeights(int n, const int* __restrict__ diag, const double* __restrict__ val, double* __restrict__ w,
                            const double* __restrict__ acc, int* __restrict__ fallbacks) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* D0 = val + (long long)diag[i] * NB * NB;
    double D[NB * NB];
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) D[e] = D0[e];
    bool inverted = false;
    if (acc) {
        const double* B = acc + (long long)i * NB * NB;
        double inv[NB * NB];
#pragma unroll
        for (int e = 0; e < NB * NB; ++e) inv[e] = B[e];
        if (invertBlockDev<NB>(inv)) {
            double n1 = 0, ni = 0;
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                double c = 0, d = 0;
#pragma unroll
                for (int r = 0; r < NB; ++r) { c += fabs(B[r * NB + j]); d += fabs(inv[r * NB + j]); }
                n1 = fmax(n1, c);
                ni = fmax(ni, d);
            }
            if (NB == 1 || n1 * ni <= 1e10) {
#pragma unroll
                for (int e = 0; e < NB * NB; ++e) D[e] = inv[e];
                inverted = true;
            } else if (fallbacks) {
                atomicAdd(fallbacks, 1);
            }
        } else if (NB == 1) {
            D[0] = B[0];
        } else if (fallbacks) {
            atomicAdd(fallbacks, 1);
        }
    }
    double wi[NB];
    if (inverted || invertBlockDev<NB>(D)) {
#pragma unroll
        for (int e = 0; e < NB; ++e) wi[e] = D[e];
    } else {
        double s = 0;
#pragma unroll
        for (int e = 0; e < NB; ++e) s += D0[e * NB];
#pragma unroll
        for (int e = 0; e < NB; ++e) wi[e] = fabs(s) > 1e-300 ? 1.0 / s : 1.0;
    }
    double wm = 0;
#pragma unroll
    for (int e = 0; e < NB; ++e) wm = fmax(wm, fabs(wi[e]));
    if (!(wm > 0) || !isfinite(wm)) {
#pragma unroll
        for (int e = 0; e < NB; ++e) wi[e] = 1.0;
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) w[(long long)i * NB + e] = wi[e];
}
// Ap_ij = sum_e w_ie S_ij[e][0] on the pattern of S
template <int NB>
__global__ void kCprPressure(int n, const int* __restrict__ rp, const double* __restrict__ val, const double* __restrict__ w,
                             double* __restrict__ ap) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double wi[NB];
#pragma unroll
    for (int e = 0; e < NB; ++e) wi[e] = w[(long long)i * NB + e];
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const double* a = val + (long long)p * NB * NB;
        double s = 0;
#pragma unroll
        for (int e = 0; e < NB; ++e) s += wi[e] * a[e * NB];
        ap[p] = s;
    }
}
__global__ void kFixDiag(int n, const int* __restrict__ diag, double* __restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double& d = v[diag[i]];
    if (fabs(d) < 1e-300) d = 1.0;
}
// r_p = w^T r per block row; with x1 given also the first damped-Jacobi iterate of level 0, x1 = omega D^-1 r_p
template <int NB, typename TD>
__global__ void kCprRestrict(int n, const double* __restrict__ w, const double* __restrict__ r, double* __restrict__ rp,
                             const TD* __restrict__ dinv, double omega, double* __restrict__ x1) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0;
#pragma unroll
    for (int e = 0; e < NB; ++e) s += w[(long long)i * NB + e] * r[(long long)i * NB + e];
    rp[i] = s;
    if (x1) x1[i] = omega * double(dinv[i]) * s;
}
// second-stage residual t = r - S(:, 0) x_p over the SpMV slices with the pressure column packed (nb values per
// slot): acc_e -= S_p[e][0] x_p,col in block order (bsr.cpp:1071-1081).  With y0 given, t of the rows whose ILU
// working index w = iperm[row] is below w1 also goes to y0[w]: the forward sweep of a first level without lower
// entries only copies t there.
template <int NB, typename TA, typename CT>
__global__ __launch_bounds__(kBlock) void kCprSecond(int nslices, const int2* __restrict__ info, const int* __restrict__ row,
                                                     const CT* __restrict__ col, const TA* __restrict__ colP, int nlong,
                                                     const int* __restrict__ lrow, const int* __restrict__ lptr,
                                                     const int* __restrict__ lcol, const double* __restrict__ lvalFull,
                                                     const double* __restrict__ r, const double* __restrict__ xp, double* __restrict__ t,
                                                     const int* __restrict__ iperm, int w1, double* __restrict__ y0) {
#pragma clang fp contract(off)
    const int wid = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wid < nslices) {
        const int rr = row[(long long)wid * kWave + lane];
        if (rr < 0) return;
        const int2 inf = info[wid];
        double acc[NB];
#pragma unroll
        for (int e = 0; e < NB; ++e) acc[e] = r[(long long)rr * NB + e];
        // four slots per round: columns and pressure-column values, then the x_p gathers, all in flight before
        // the first product; the updates are applied slot by slot as before
        int k = 0;
        for (; k + 4 <= inf.x; k += 4) {
            const long long slot = (long long)inf.y + k;
            double av[4][NB], xk[4];
            int c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                c[u] = slotCol<CT, true>(col, (slot + u) * kWave + lane, rr);
                const TA* a = colP + (slot + u) * (NB * kWave) + lane;
#pragma unroll
                for (int e = 0; e < NB; ++e) av[u][e] = double(__builtin_nontemporal_load(a + e * kWave));
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) xk[u] = xp[c[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int e = 0; e < NB; ++e) acc[e] -= av[u][e] * xk[u];
        }
        for (; k < inf.x; ++k) {
            const long long slot = (long long)inf.y + k;
            const double xk = xp[slotCol<CT, true>(col, slot * kWave + lane, rr)];
            const TA* a = colP + slot * (NB * kWave) + lane;
#pragma unroll
            for (int e = 0; e < NB; ++e) acc[e] -= double(__builtin_nontemporal_load(a + e * kWave)) * xk;
        }
#pragma unroll
        for (int e = 0; e < NB; ++e) t[(long long)rr * NB + e] = acc[e];
        if (y0) {
            const int w = iperm[rr];
            if (w < w1) {
#pragma unroll
                for (int e = 0; e < NB; ++e) y0[(long long)w * NB + e] = acc[e];
            }
        }
        return;
    }
    const int q = wid - nslices;
    if (q >= nlong) return;
    double part[NB];
#pragma unroll
    for (int e = 0; e < NB; ++e) part[e] = 0;
    for (int p = lptr[q] + lane; p < lptr[q + 1]; p += kWave) {
        const double xk = xp[lcol[p]];
        const double* a = lvalFull + (long long)p * NB * NB;
#pragma unroll
        for (int e = 0; e < NB; ++e) part[e] += a[e * NB] * xk;
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) part[e] = waveSum(part[e]);
    if (lane == 0) {
        const int rr = lrow[q];
        const int w = y0 ? iperm[rr] : w1;
#pragma unroll
        for (int e = 0; e < NB; ++e) {
            const double v = r[(long long)rr * NB + e] - part[e];
            t[(long long)rr * NB + e] = v;
            if (w < w1) y0[(long long)w * NB + e] = v;
        }
    }
}

// ---------------------------------------------------------------- scalar AMG level kernels (amg.cpp:98-131, 303-329)
__global__ void kCsrResidual(int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
                             const double* __restrict__ x, const double* __restrict__ b, double* __restrict__ r) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) s += v[p] * x[ci[p]];
    r[i] = b ? b[i] - s : s;
}
// One level of a Gauss-Seidel sweep.  Forward: columns below i read xout (this sweep), above i read xin (before
// the sweep, or zero when xin is null); backward the other way round.  s = b_i - sum_{p != diag} a_p x_col in
// column order, x_i = s / a_ii (amg.cpp:98-113).
__global__ void kGsLevel(int nrows, const int* __restrict__ rows, const int* __restrict__ rp, const int* __restrict__ ci,
                         const int* __restrict__ diag, const double* __restrict__ v, const double* __restrict__ b,
                         const double* __restrict__ xin, double* __restrict__ xout, int forward) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    const int i = rows[t];
    double s = b[i];
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const int j = ci[p];
        if (j == i) continue;
        const bool useOut = forward ? (j < i) : (j > i);
        if (useOut) s -= v[p] * xout[j];
        else if (xin) s -= v[p] * xin[j];
    }
    xout[i] = s / v[diag[i]];
}
// restriction: the sum of the fine residuals of each aggregate, G lanes per aggregate; with x1 given also the
// coarse level's first damped-Jacobi iterate x1 = omega D^-1 b_c
template <int G, typename TD>
__global__ __launch_bounds__(256) void kRestrict(int nc, const int* __restrict__ rrp, const int* __restrict__ rci, const double* __restrict__ r,
                                                 double* __restrict__ bc, const TD* __restrict__ dinv, double omega, double* __restrict__ x1) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int a = int(t / G), sub = int(t % G);
    if (a >= nc) return;
    double s = 0;
    for (int q = rrp[a] + sub; q < rrp[a + 1]; q += G) s += r[rci[q]];
    s = groupSum<G>(s);
    if (sub == 0) {
        bc[a] = s;
        if (x1) x1[a] = omega * double(dinv[a]) * s;
    }
}
// x1 = omega D^-1 b (the first damped-Jacobi iterate from zero)
template <typename TD>
__global__ void kJacobiInit(int n, const TD* __restrict__ dinv, const double* __restrict__ b, double omega, double* __restrict__ x1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x1[i] = omega * double(dinv[i]) * b[i];
}
__global__ void kProlongAdd(int n, const int* __restrict__ agg, const double* __restrict__ xc, double* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = x[i] + xc[agg[i]];
}
// y = M x for a dense m x ld row-major matrix (the explicit coarsest inverse): one wavefront per row; with pfine
// given, also the prolongation into the finer level's iterate (pfine at the members prp / pci of aggregate row)
__global__ __launch_bounds__(kBlock) void kDenseGemv(int m, int ld, const double* __restrict__ M, const double* __restrict__ x,
                                                     double* __restrict__ y, const int* __restrict__ prp, const int* __restrict__ pci,
                                                     double* __restrict__ pfine) {
    const int rowi = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (rowi >= m) return;
    // exact zeros of the inverse are skipped: a batch's coarsest matrix is block-diagonal over its systems, and a
    // non-finite value of one system must not reach the others through 0 * x
    double s = 0;
    for (int j = lane; j < m; j += kWave) {
        const double a = M[(long long)rowi * ld + j];
        if (a != 0.0) s += a * x[j];
    }
    s = waveSum(s);                            // every lane holds the sum
    if (lane == 0) y[rowi] = s;
    if (pfine)
        for (int q = prp[rowi] + lane; q < prp[rowi + 1]; q += kWave) {
            const int j = pci[q];
            pfine[j] = pfine[j] + s;
        }
}
__global__ void kGalerkinSum(int count, const int* __restrict__ ptr, const int* __restrict__ srcIdx, const double* __restrict__ src,
                             double* __restrict__ dst) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const int u0 = ptr[q], u1 = ptr[q + 1];
    double s = src[srcIdx[u0]];
    for (int u = u0 + 1; u < u1; ++u) s += src[srcIdx[u]];
    dst[q] = s;
}
// [C | I] (m x 2m) from the coarsest CSR
__global__ void kDenseFromCsr(int m, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
                              double* __restrict__ M) {
    const long long total = (long long)m * 2 * m;
    for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total; idx += (long long)gridDim.x * blockDim.x) {
        const int i = int(idx / (2 * m)), j = int(idx % (2 * m));
        double val = 0;
        if (j < m) {
            for (int p = rp[i]; p < rp[i + 1]; ++p) if (ci[p] == j) val += v[p];
        } else if (j - m == i) val = 1.0;
        M[idx] = val;
    }
}
// One Gauss-Jordan step with partial pivoting on [C | I], out of place (Min -> Mout).  Every workgroup finds the
// pivot row of column k itself (first maximum), so the step needs no grid-wide synchronisation.
__global__ __launch_bounds__(kBlock) void kGaussJordanStep(int m, int k, const double* __restrict__ Min, double* __restrict__ Mout) {
    __shared__ double sv[kBlock];
    __shared__ int si[kBlock];
    const int ld = 2 * m;
    double best = -1;
    int bi = k;
    for (int i = k + threadIdx.x; i < m; i += kBlock) {
        const double a = fabs(Min[(long long)i * ld + k]);
        if (a > best) { best = a; bi = i; }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const double a = sv[threadIdx.x + o];
            const int ia = si[threadIdx.x + o];
            if (a > sv[threadIdx.x] || (a == sv[threadIdx.x] && ia < si[threadIdx.x])) { sv[threadIdx.x] = a; si[threadIdx.x] = ia; }
        }
        __syncthreads();
    }
    const int pr = si[0];
    double piv = Min[(long long)pr * ld + k];
    if (fabs(piv) < 1e-300) piv = piv < 0 ? -1e-300 : 1e-300;
    const long long total = (long long)m * ld;
    for (long long idx = blockIdx.x * (long long)kBlock + threadIdx.x; idx < total; idx += (long long)gridDim.x * kBlock) {
        const int i = int(idx / ld), j = int(idx % ld);
        const double pk = Min[(long long)pr * ld + j] / piv;    // new row k
        double out;
        if (i == k) out = pk;
        else {
            const int src = i == pr ? k : i;                     // row pr receives the old row k
            const double f = Min[(long long)src * ld + k];
            out = Min[(long long)src * ld + j] - f * pk;
        }
        Mout[idx] = out;
    }
}

// argmax |v_i| over the rows a thread was handed (first index on ties), then over the workgroup (LDS tree, first
// index on ties): the result does not depend on how the rows were split
__device__ inline int gjArgmax(double best, int bi, double* sv, int* si) {
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int o = kBlock / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            const double a = sv[threadIdx.x + o];
            const int ia = si[threadIdx.x + o];
            if (a > sv[threadIdx.x] || (a == sv[threadIdx.x] && ia < si[threadIdx.x])) { sv[threadIdx.x] = a; si[threadIdx.x] = ia; }
        }
        __syncthreads();
    }
    const int r = si[0];
    __syncthreads();
    return r;
}
// Two Gauss-Jordan steps (k, k + 1) in one launch, bit-identical to two kGaussJordanStep launches: every value of
// the intermediate matrix B (after step k) that step k + 1 needs is recomputed from Min with the same expressions
// (B_rc = Min[pr][c] / piv for r = k, Min[src(r)][c] - Min[src(r)][k] * (Min[pr][c] / piv) otherwise), the pivot
// of step k + 1 is searched on B's column k + 1 with the same first-index rule.
__global__ __launch_bounds__(kBlock) void kGaussJordanStep2(int m, int k, const double* __restrict__ Min, double* __restrict__ Mout) {
    __shared__ double sv[kBlock];
    __shared__ int si[kBlock];
    const int ld = 2 * m;
    double best = -1;
    int bi = k;
    for (int i = k + threadIdx.x; i < m; i += kBlock) {
        const double a = fabs(Min[(long long)i * ld + k]);
        if (a > best) { best = a; bi = i; }
    }
    const int pr1 = gjArgmax(best, bi, sv, si);
    double piv1 = Min[(long long)pr1 * ld + k];
    if (fabs(piv1) < 1e-300) piv1 = piv1 < 0 ? -1e-300 : 1e-300;
    auto b1 = [&](int r, int c) -> double {                   // the matrix after step k
        const double pk = Min[(long long)pr1 * ld + c] / piv1;
        if (r == k) return pk;
        const int src = r == pr1 ? k : r;
        const double f = Min[(long long)src * ld + k];
        return Min[(long long)src * ld + c] - f * pk;
    };
    const int k2 = k + 1;
    best = -1;
    bi = k2;
    for (int i = k2 + threadIdx.x; i < m; i += kBlock) {
        const double a = fabs(b1(i, k2));
        if (a > best) { best = a; bi = i; }
    }
    const int pr2 = gjArgmax(best, bi, sv, si);
    double piv2 = b1(pr2, k2);
    if (fabs(piv2) < 1e-300) piv2 = piv2 < 0 ? -1e-300 : 1e-300;
    const long long total = (long long)m * ld;
    for (long long idx = blockIdx.x * (long long)kBlock + threadIdx.x; idx < total; idx += (long long)gridDim.x * kBlock) {
        const int i = int(idx / ld), j = int(idx % ld);
        const double pk = b1(pr2, j) / piv2;
        double out;
        if (i == k2) out = pk;
        else {
            const int src = i == pr2 ? k2 : i;
            const double f = b1(src, k2);
            out = b1(src, j) - f * pk;
        }
        Mout[idx] = out;
    }
}

// ---------------------------------------------------------------- fused damped-Jacobi V-cycle steps (DeviceCprAmg smoothing)
// The first iterate from zero, x1 = w D^-1 b, is written by the restriction that produces b, and the prolongation
// x += P x_c by the coarse level's last kernel, so every step gathers one vector per entry:
// MODE 0 (residual): out0 = b - A x1.
// MODE 1 (smoothing step): out0 = x1 + w D^-1 (b - A x1).
// the row's own operands do not depend on the row's entries: their loads are issued before the entry loop so that
// they overlap the gathers
struct JacSelf { double d, b, x; };
template <int MODE, typename TD>
__device__ inline JacSelf jacSelf(int i, const TD* __restrict__ dinv, const double* __restrict__ b, const double* __restrict__ x1) {
    JacSelf o;
    o.d = MODE == 0 ? 0.0 : double(dinv[i]);
    o.b = b[i];
    o.x = MODE == 0 ? 0.0 : x1[i];
    return o;
}
template <int MODE>
__device__ inline double jacValue(double acc, double omega, const JacSelf& o) {
    if (MODE == 0) return o.b - acc;
    return o.x + omega * o.d * (o.b - acc);
}
// level 0 over the SpMV row slices (the pressure matrix has the pattern of S), long rows one wavefront each
template <int MODE, typename TA, typename CT>
__global__ __launch_bounds__(kBlock) void kJacobiSliced(int nslices, const int2* __restrict__ info, const int* __restrict__ row,
                                                        const CT* __restrict__ col, const TA* __restrict__ val, int nlong,
                                                        const int* __restrict__ lrow, const int* __restrict__ lptr, const int* __restrict__ lcol,
                                                        const TA* __restrict__ lval, const TA* __restrict__ dinv,
                                                        const double* __restrict__ b, const double* __restrict__ x1, double omega,
                                                        double* __restrict__ out0) {
    const int wid = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wid < nslices) {
        const int i = row[(long long)wid * kWave + lane];
        if (i < 0) return;
        const JacSelf self = jacSelf<MODE>(i, dinv, b, x1);
        const int2 inf = info[wid];
        double acc = 0;
        // four slots per round: their column and value loads, then their x gathers, are all in flight before the
        // first product (the products are still added one by one in slot order)
        int k = 0;
        for (; k + 4 <= inf.x; k += 4) {
            const long long at = ((long long)inf.y + k) * kWave + lane;
            double v[4], xv[4];
            int c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v[u] = double(val[at + u * kWave]);
                c[u] = slotCol<CT, false>(col, at + u * kWave, i);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) xv[u] = x1[c[u]];
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += v[u] * xv[u];
        }
        for (; k < inf.x; ++k) {
            const long long at = ((long long)inf.y + k) * kWave + lane;
            acc += double(val[at]) * x1[slotCol<CT, false>(col, at, i)];
        }
        out0[i] = jacValue<MODE>(acc, omega, self);
        return;
    }
    const int q = wid - nslices;
    if (q >= nlong) return;
    const int i = lrow[q];
    JacSelf self{};
    if (lane == 0) self = jacSelf<MODE>(i, dinv, b, x1);
    double acc = 0;
    for (int p = lptr[q] + lane; p < lptr[q + 1]; p += kWave) acc += double(lval[p]) * x1[lcol[p]];
    acc = waveSum(acc);
    if (lane == 0) out0[i] = jacValue<MODE>(acc, omega, self);
}
// coarse levels in row slices (one row per lane), with the arithmetic of the G-lane CSR step below: entry k of a row
// goes to chain k mod G (the sum lane k mod G forms there), and the G chains are folded in the order of the
// groupSum butterfly (offsets G/2, G/4, ..., 1), so the result is bit-identical to kJacobiCsr<MODE, G>.  Padding
// slots add 0 * x_own.  Four slots per round are in flight before their products; long rows take G lanes each as
// in kJacobiCsr (64 / G of them per wavefront).
template <int MODE, int G, typename TA, typename CT>
__global__ __launch_bounds__(kBlock) void kJacobiSlicedG(int nslices, const int2* __restrict__ info, const int* __restrict__ row,
                                                         const CT* __restrict__ col, const TA* __restrict__ val, int nlong,
                                                         const int* __restrict__ lrow, const int* __restrict__ lptr, const int* __restrict__ lcol,
                                                         const TA* __restrict__ lval, const TA* __restrict__ dinv,
                                                         const double* __restrict__ b, const double* __restrict__ x1, double omega,
                                                         double* __restrict__ out0) {
    static_assert(G % 4 == 0 && G <= kWave, "chains in rounds of four slots");
    const int wid = blockIdx.x * (kBlock / kWave) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (wid < nslices) {
        const int i = row[(long long)wid * kWave + lane];
        if (i < 0) return;
        const JacSelf self = jacSelf<MODE>(i, dinv, b, x1);
        const int2 inf = info[wid];
        double a[G];
#pragma unroll
        for (int j = 0; j < G; ++j) a[j] = 0;
        for (int kb = 0; kb < inf.x; kb += G) {
#pragma unroll
            for (int r = 0; r < G; r += 4) {
                if (kb + r >= inf.x) break;
                const long long at = ((long long)inf.y + kb + r) * kWave + lane;
                double v[4], xv[4];
                int c[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const bool in = kb + r + u < inf.x;
                    v[u] = in ? double(val[at + u * kWave]) : 0.0;
                    c[u] = in ? slotCol<CT, false>(col, at + u * kWave, i) : i;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) xv[u] = x1[c[u]];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (kb + r + u < inf.x) a[r + u] += v[u] * xv[u];
            }
        }
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
            for (int j = 0; j < o; ++j) a[j] = a[j] + a[j + o];
        out0[i] = jacValue<MODE>(a[0], omega, self);
        return;
    }
    const int q = (wid - nslices) * (kWave / G) + lane / G, sub = lane % G;
    if (q >= nlong) return;                  // whole groups leave together
    const int i = lrow[q];
    const JacSelf self = jacSelf<MODE>(i, dinv, b, x1);
    double acc = 0;
    for (int p = lptr[q] + sub; p < lptr[q + 1]; p += G) acc += double(lval[p]) * x1[lcol[p]];
    acc = groupSum<G>(acc);
    if (sub == 0) out0[i] = jacValue<MODE>(acc, omega, self);
}
// coarse levels (CSR): G lanes per row (G a power of two near the mean row length), the row's entries strided over
// the group so their loads are issued together, then a shuffle reduction in a fixed order.  With pfine given (the
// last post-smoothing step of a coarse level) the group also adds the new value to the finer level's iterate at
// the members of its aggregate (prp / pci): the prolongation x_f += P x_c, every fine row written by one group.
template <int MODE, int G, typename TA>
__global__ __launch_bounds__(256) void kJacobiCsr(int n, const int* __restrict__ rp, const int* __restrict__ ci, const TA* __restrict__ v,
                                                  const TA* __restrict__ dinv, const double* __restrict__ b, const double* __restrict__ x1,
                                                  double omega, double* __restrict__ out0, const int* __restrict__ prp,
                                                  const int* __restrict__ pci, double* __restrict__ pfine) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = int(t / G), sub = int(t % G);
    if (i >= n) return;                       // whole groups leave together (n * G is a multiple of G)
    const JacSelf self = jacSelf<MODE>(i, dinv, b, x1);
    double acc = 0;
    for (int p = rp[i] + sub; p < rp[i + 1]; p += G) acc += double(v[p]) * x1[ci[p]];
    acc = groupSum<G>(acc);                   // every lane of the group holds the sum
    const double val = jacValue<MODE>(acc, omega, self);
    if (sub == 0) out0[i] = val;
    if (pfine)
        for (int q = prp[i] + sub; q < prp[i + 1]; q += G) {
            const int j = pci[q];
            pfine[j] = pfine[j] + val;
        }
}
__global__ void kDinv(int n, const int* __restrict__ diag, const double* __restrict__ v, double* __restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = v[diag[i]];
    dinv[i] = fabs(d) > 1e-300 ? 1.0 / d : 0.0;
}

// ---------------------------------------------------------------- vectors and deterministic reductions
__global__ void kFill(double* __restrict__ x, double v, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) x[i] = v;
}
__global__ void kCopy(double* __restrict__ d, const double* __restrict__ s, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) d[i] = s[i];
}
__global__ void kCopyVec(const double2* __restrict__ s, double2* __restrict__ d, long long n2) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n2; i += (long long)gridDim.x * blockDim.x) d[i] = s[i];
}

// Every state but kRun freezes the scalars, so iterations launched after the solve stopped change nothing: the
// early exit on ||s|| (kConvS) lets the x update add alpha ph once, and the closing reduction of that iteration
// (kEpRn / kEpRnRho) freezes it (kDoneS) before the next iteration's update could run again.
__device__ void runEpilogue(int ep, const double* d, KScalars* ks) {
    if (!ks) return;
    if (ks->state == kConvS) {
        if (ep == kEpRn || ep == kEpRnRho) ks->state = kDoneS;
        return;
    }
    if (ks->state != kRun) return;
    switch (ep) {
        case kEpRho:
            ks->rhoN = d[0];
            if (fabs(d[0]) < 1e-300) { ks->state = kBreak; break; }
            ks->beta = (d[0] / ks->rho) * (ks->alpha / ks->omega);
            ks->rho = d[0];
            break;
        case kEpAlpha:
            ks->rv = d[0];
            if (fabs(d[0]) < 1e-300) { ks->state = kBreak; break; }
            ks->alpha = ks->rho / d[0];
            break;
        case kEpSn:
            ks->sn = sqrt(d[0]);
            if (ks->sn / ks->bn < ks->tol) { ks->state = kConvS; ks->it += 1; }
            break;
        case kEpOmega:
            ks->tt = d[0];
            ks->ts = d[1];
            ks->omega = d[0] > 0 ? d[1] / d[0] : 0;
            break;
        case kEpRn:
            ks->rn = sqrt(d[0]);
            ks->it += 1;
            if (ks->rn / ks->bn < ks->tol) ks->state = kConv;
            else if (ks->omega == 0) ks->state = kOmegaZero;
            break;
        case kEpRnRho:   // kEpRn, then the next iteration's kEpRho with d[1] = (rh, r)
            ks->rn = sqrt(d[0]);
            ks->it += 1;
            if (ks->rn / ks->bn < ks->tol) { ks->state = kConv; break; }
            if (ks->omega == 0) { ks->state = kOmegaZero; break; }
            ks->rhoN = d[1];
            if (fabs(d[1]) < 1e-300) { ks->state = kBreak; break; }
            ks->beta = (d[1] / ks->rho) * (ks->alpha / ks->omega);
            ks->rho = d[1];
            break;
        case kEpBn: ks->bn = sqrt(d[0]); break;
        default: break;
    }
}

// Deterministic two-launch reductions: every block writes its K partial sums (fixed tree inside the block), then
// one block sums the partials in block order and runs the epilogue.  The kernel boundary publishes the partials,
// so no block needs a device-wide fence (on this multi-XCD part such a fence writes back the whole L2).
template <int K>
__device__ void reducePartial(double (&v)[K], double* __restrict__ partial) {
    __shared__ double sh[K][kBlock / kWave];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double t = waveSum(v[k]);
        if (lane == 0) sh[k][w] = t;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0;
        for (int q = 0; q < kBlock / kWave; ++q) s += sh[threadIdx.x][q];
        partial[(long long)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}
__global__ __launch_bounds__(kBlock) void kReduceFinal(int K, int G, const double* __restrict__ partial, double* out, int ep, KScalars* ks) {
    __shared__ double sh[32][kBlock / kWave];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < K; ++k) {
        double acc = 0;
        for (int b = threadIdx.x; b < G; b += kBlock) acc += partial[(long long)k * G + b];
        acc = waveSum(acc);
        if (lane == 0) sh[k][w] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double res[32];
        for (int k = 0; k < K; ++k) {
            double s = 0;
            for (int q = 0; q < kBlock / kWave; ++q) s += sh[k][q];
            res[k] = s;
            if (out) out[k] = s;
        }
        runEpilogue(ep, res, ks);
    }
}

template <int K>
__global__ __launch_bounds__(kBlock) void kDots(long long n, const double* __restrict__ a0, const double* __restrict__ b0,
                                                const double* __restrict__ a1, const double* __restrict__ b1, double* partial) {
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        v[0] += a0[i] * b0[i];
        if constexpr (K > 1) v[1] += a1[i] * b1[i];
    }
    reducePartial<K>(v, partial);
}

// p = r + beta (p - omega v)
__global__ void kBicgP(long long n, const KScalars* __restrict__ ks, const double* __restrict__ r, const double* __restrict__ v,
                       double* __restrict__ p) {
#pragma clang fp contract(off)
    if (ks->state != kRun) return;
    const double beta = ks->beta, omega = ks->omega;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = r[i] + beta * (p[i] - omega * v[i]);
}
// s = r - alpha v, with (s, s)
__global__ __launch_bounds__(kBlock) void kBicgS(long long n, const KScalars* ks, const double* __restrict__ r, const double* __restrict__ v,
                                                 double* __restrict__ sv, double* partial) {
#pragma clang fp contract(off)
    double acc[1] = {0};
    if (ks->state == kRun) {
        const double alpha = ks->alpha;
        for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
            const double s = r[i] - alpha * v[i];
            sv[i] = s;
            acc[0] += s * s;
        }
    }
    reducePartial<1>(acc, partial);
}
// x += alpha ph + omega sh; r = s - omega t; with (r, r) and the next iteration's (rh, r) (accumulated with a fused
// multiply-add, as kDots accumulates it, so the value is the one kDots gives); after the early exit on ||s||
// (kConvS) only x += alpha ph
__global__ __launch_bounds__(kBlock) void kBicgXR(long long n, const KScalars* ks, const double* __restrict__ ph, const double* __restrict__ sh,
                                                  const double* __restrict__ sv, const double* __restrict__ t, const double* __restrict__ rh,
                                                  double* __restrict__ x, double* __restrict__ r, double* partial) {
#pragma clang fp contract(off)
    double acc[2] = {0, 0};
    const int state = ks->state;
    if (state == kRun) {
        const double alpha = ks->alpha, omega = ks->omega;
        for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
            x[i] += alpha * ph[i] + omega * sh[i];
            const double ri = sv[i] - omega * t[i];
            r[i] = ri;
            acc[0] += ri * ri;
            acc[1] = fma(rh[i], ri, acc[1]);
        }
    } else if (state == kConvS) {
        const double alpha = ks->alpha;
        for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) x[i] += alpha * ph[i];
    }
    reducePartial<2>(acc, partial);
}
// ---------------------------------------------------------------- batched BiCGSTAB (independent systems, per-system scalars)
// The flat vectors hold every system; task t = {system, first entry, end entry} is one workgroup of the vector
// kernels, so each workgroup reads its system's scalars once.  Reductions: one partial per task (fixed tree), then
// one thread per system adds its tasks' partials in task order and runs the system's epilogue.
__device__ inline void batchEpilogue(int ep, const double* d, KScalars* ks) {
    if (ep == kEpBn && ks->state == kRun) {
        ks->bn = sqrt(d[0]);
        if (ks->bn == 0) ks->state = kConv;   // zero right-hand side: x = 0
        return;
    }
    runEpilogue(ep, d, ks);
}
template <int K>
__device__ void taskPartial(double (&v)[K], double* __restrict__ partial, int T) {
    __shared__ double sh[K][kBlock / kWave];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double t = waveSum(v[k]);
        if (lane == 0) sh[k][w] = t;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        double s = 0;
        for (int q = 0; q < kBlock / kWave; ++q) s += sh[threadIdx.x][q];
        partial[(long long)threadIdx.x * T + blockIdx.x] = s;
    }
}
__global__ void kBFinal(int ns, int K, const int* __restrict__ tStart, const double* __restrict__ partial, int T, int ep, KScalars* ks,
                        double* __restrict__ out) {
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= ns) return;
    double d[2] = {0, 0};
    for (int k = 0; k < K; ++k)
        for (int t = tStart[sys]; t < tStart[sys + 1]; ++t) d[k] += partial[(long long)k * T + t];
    if (out)
        for (int k = 0; k < K; ++k) out[(long long)sys * K + k] = d[k];
    if (ks) batchEpilogue(ep, d, ks + sys);
}
// zero the entries of the systems whose flag is 0
__global__ __launch_bounds__(kBlock) void kBMask(const int4* __restrict__ task, const int* __restrict__ keep, double* __restrict__ v) {
    const int4 tk = task[blockIdx.x];
    if (keep[tk.x]) return;
    for (int i = tk.y + threadIdx.x; i < tk.z; i += kBlock) v[i] = 0.0;
}
// x += d over every task (refinement step)
__global__ __launch_bounds__(kBlock) void kBAdd(const int4* __restrict__ task, const double* __restrict__ d, double* __restrict__ x) {
    const int4 tk = task[blockIdx.x];
    for (int i = tk.y + threadIdx.x; i < tk.z; i += kBlock) x[i] += d[i];
}
template <int K>
__global__ __launch_bounds__(kBlock) void kBDots(const int4* __restrict__ task, int T, const double* __restrict__ a0, const double* __restrict__ b0,
                                                 const double* __restrict__ a1, const double* __restrict__ b1, double* partial) {
    const int4 tk = task[blockIdx.x];
    double v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = 0;
    for (int i = tk.y + threadIdx.x; i < tk.z; i += kBlock) {
        v[0] += a0[i] * b0[i];
        if constexpr (K > 1) v[1] += a1[i] * b1[i];
    }
    taskPartial<K>(v, partial, T);
}
__global__ __launch_bounds__(kBlock) void kBBicgP(const int4* __restrict__ task, const KScalars* __restrict__ ks, const double* __restrict__ r,
                                                  const double* __restrict__ v, double* __restrict__ p) {
#pragma clang fp contract(off)
    const int4 tk = task[blockIdx.x];
    const KScalars& k = ks[tk.x];
    if (k.state != kRun) return;
    const double beta = k.beta, omega = k.omega;
    for (int i = tk.y + threadIdx.x; i < tk.z; i += kBlock) p[i] = r[i] + beta * (p[i] - omega * v[i]);
}
__global__ __launch_bounds__(kBlock) void kBBicgS(const int4* __restrict__ task, int T, const KScalars* __restrict__ ks, const double* __restrict__ r,
                                                  const double* __restrict__ v, double* __restrict__ sv, double* partial) {
#pragma clang fp contract(off)
    const int4 tk = task[blockIdx.x];
    const KScalars& k = ks[tk.x];
    double acc[1] = {0};
    if (k.state == kRun) {
        const double alpha = k.alpha;
        for (int i = tk.y + threadIdx.x; i < tk.z; i += kBlock) {
            const double s = r[i] - alpha * v[i];
            sv[i] = s;
            acc[0] += s * s;
        }
    }
    taskPartial<1>(acc, partial, T);
}
__global__ __launch_bounds__(kBlock) void kBBicgXs(const int4* __restrict__ task, const KScalars* __restrict__ ks, const double* __restrict__ ph,
                                                   double* __restrict__ x) {
#pragma clang fp contract(off)
    const int4 tk = task[blockIdx.x];
    const KScalars& k = ks[tk.x];
    if (k.state != kConvS) return;
    const double alpha = k.alpha;
    for (int i = tk.y + threadIdx.x; i < tk.z; i += kBlock) x[i] += alpha * ph[i];
}
__global__ __launch_bounds__(kBlock) void kBBicgXR(const int4* __restrict__ task, int T, const KScalars* __restrict__ ks, const double* __restrict__ ph,
                                                   const double* __restrict__ sh, const double* __restrict__ sv, const double* __restrict__ t,
                                                   double* __restrict__ x, double* __restrict__ r, double* partial) {
#pragma clang fp contract(off)
    const int4 tk = task[blockIdx.x];
    const KScalars& k = ks[tk.x];
    double acc[1] = {0};
    if (k.state == kRun) {
        const double alpha = k.alpha, omega = k.omega;
        for (int i = tk.y + threadIdx.x; i < tk.z; i += kBlock) {
            x[i] += alpha * ph[i] + omega * sh[i];
            const double ri = sv[i] - omega * t[i];
            r[i] = ri;
            acc[0] += ri * ri;
        }
    }
    taskPartial<1>(acc, partial, T);
}
__global__ void kBInit(int ns, const double* __restrict__ tol, KScalars* ks) {
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= ns) return;
    KScalars k{};
    k.rho = 1; k.alpha = 1; k.omega = 1; k.tol = tol[sys]; k.state = kRun; k.it = 0;
    ks[sys] = k;
}

__global__ void kAxpy(long long n, double a, const double* __restrict__ x, double* __restrict__ y) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) y[i] += a * x[i];
}
__global__ void kAxpyDev(long long n, const double* __restrict__ a, double sign, const double* __restrict__ x, double* __restrict__ y) {
    const double aa = sign * a[0];
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) y[i] += aa * x[i];
}
__global__ void kScaleTo(long long n, const double* __restrict__ num, const double* __restrict__ x, double* __restrict__ y, int invert) {
    const double a = num[0];
    const double f = invert ? (a > 0 ? 1.0 / a : 0.0) : a;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) y[i] = x[i] * f;
}
__global__ void kSub(long long n, const double* __restrict__ b, const double* __restrict__ ax, double* __restrict__ r) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) r[i] = b[i] - ax[i];
}
// partial sums of (V_i, w), i < k <= 32
__global__ __launch_bounds__(kBlock) void kMultiDot(long long n, int k, const double* __restrict__ V, long long ld,
                                                    const double* __restrict__ w, double* partial) {
    __shared__ double sh[32][kBlock / kWave];
    double acc[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) acc[q] = 0;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        const double wi = w[i];
#pragma unroll
        for (int q = 0; q < 32; ++q) if (q < k) acc[q] += V[q * ld + i] * wi;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < k; ++q) {
        const double t = waveSum(acc[q]);
        if (lane == 0) sh[q][wv] = t;
    }
    __syncthreads();
    if (threadIdx.x < k) {
        double s = 0;
        for (int u = 0; u < kBlock / kWave; ++u) s += sh[threadIdx.x][u];
        partial[(long long)threadIdx.x * gridDim.x + blockIdx.x] = s;
    }
}
__global__ void kMultiAxpy(long long n, int k, const double* __restrict__ V, long long ld, const double* __restrict__ h,
                           double* __restrict__ w, double sign) {
    __shared__ double hs[32];
    if (threadIdx.x < k) hs[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double s = 0;
        for (int q = 0; q < k; ++q) s += hs[q] * V[q * ld + i];
        w[i] += sign * s;
    }
}

template <int NB, typename TA, int LN>
void spmvLanesT(hipStream_t s, const SliceDev<TA>& d, const double* x, double* y, const KScalars* gate) {
    const dim3 grid((d.nslices + d.nlong + 3) / 4);
    if (d.dcol.n)
        hipLaunchKernelGGL((kSpmvSliced<NB, TA, short, LN>), grid, dim3(kBlock), 0, s, d.nslices, d.info.p, d.row.p, d.dcol.p, d.val.p, d.nlong,
                           d.lrow.p, d.lptr.p, d.lcol.p, d.lval.p, x, y, d.owner.p, d.lowner.p, gate);
    else
        hipLaunchKernelGGL((kSpmvSliced<NB, TA, int, LN>), grid, dim3(kBlock), 0, s, d.nslices, d.info.p, d.row.p, d.col.p, d.val.p, d.nlong,
                           d.lrow.p, d.lptr.p, d.lcol.p, d.lval.p, x, y, d.owner.p, d.lowner.p, gate);
}
template <int NB, typename TA>
void spmvLayoutT(hipStream_t s, const SliceDev<TA>& d, const double* x, double* y, const KScalars* gate) {
    if (!(d.nslices + d.nlong)) return;
    if (d.lanes == 4) spmvLanesT<NB, TA, 4>(s, d, x, y, gate);
    else if (d.lanes == 2) spmvLanesT<NB, TA, 2>(s, d, x, y, gate);
    else if (d.lanes == 1) spmvLanesT<NB, TA, 1>(s, d, x, y, gate);
    else throw Error(-1, "SpMV: lanes per row must be 1, 2 or 4");
}
template <int NB>
void spmvT(hipStream_t s, bool f32, const SliceDev<double>* d64, const SliceDev<float>* d32, const double* x, double* y, const KScalars* gate) {
    if (f32) spmvLayoutT<NB, float>(s, *d32, x, y, gate);
    else spmvLayoutT<NB, double>(s, *d64, x, y, gate);
}

template <int NB, typename TA, typename TV>
void iluFwdT(hipStream_t s, const SliceDev<TA>& d, int s0, int s1, int q0, int q1, const double* r, TV* y, const KScalars* gate, double* zFused,
             const double* xp) {
    const int ns = s1 - s0, nq = q1 - q0, waves = ns + nq;
    if (!waves) return;
    const bool fused = zFused && d.dsrc.n && d.dval.n;
    hipLaunchKernelGGL((kIluForward<NB, TA, TV>), dim3((waves + 3) / 4), dim3(kBlock), 0, s, s0, ns, q0, nq, d.info.p, d.row.p, d.orig.p,
                       d.col.p, d.val.p, d.lrow.p, d.lorig.p, d.lptr.p, d.lcol.p, d.lval.p, r, y, d.owner.p, d.lowner.p, gate,
                       fused ? d.dsrc.p : nullptr, fused ? d.dval.p : nullptr, zFused, xp);
}
template <int NB, typename TA, typename TV>
void iluBwdT(hipStream_t s, const SliceDev<TA>& d, int s0, int s1, int q0, int q1, TV* y, double* z, const double* xp, const KScalars* gate,
             bool writeY) {
    const int ns = s1 - s0, nq = q1 - q0, waves = ns + nq;
    if (!waves) return;
    hipLaunchKernelGGL((kIluBackward<NB, TA, TV>), dim3((waves + 3) / 4), dim3(kBlock), 0, s, s0, ns, q0, nq, d.info.p, d.row.p,
                       d.orig.p, d.col.p, d.val.p, d.dval.p, d.lrow.p, d.lorig.p, d.lptr.p, d.lcol.p, d.lval.p, d.ldval.p, y, z, xp, d.owner.p,
                       d.lowner.p, gate, writeY ? 1 : 0);
}

template <int NB>
void bilu0LevelT(hipStream_t s, const int* rows, int r0, int r1, const int* rp, const int* ci, const int* diag, double* val, int* nreg, int stop,
                 int lean) {
    const int cnt = r1 - r0;
    hipLaunchKernelGGL(kBilu0Level<NB>, dim3((cnt + 127) / 128), dim3(128), 0, s, rows, r0, r1, rp, ci, diag, val, nreg, stop, lean);
}
template <int NB>
void iluGroupFactorT(hipStream_t s, int g0, int g1, const int* gstart, const int* grows, const long long* gidxOff, const int* gidx,
                     const int* diag, double* val, int* nreg, const int* rp, int lean) {
    hipLaunchKernelGGL(kIluGroupFactor<NB>, dim3(g1 - g0), dim3(kGroupThreads), 0, s, gstart + g0, grows, gidxOff + g0, gidx, diag, val, nreg,
                       rp, lean);
}
template <int NB>
void cprWeightsT(hipStream_t s, int n, const int* diag, const double* val, double* w, const double* acc, int* fallbacks) {
    hipLaunchKernelGGL(kCprWeights<NB>, dim3((n + 255) / 256), dim3(256), 0, s, n, diag, val, w, acc, fallbacks);
}
template <int NB>
void cprPressureT(hipStream_t s, int n, const int* rp, const double* val, const double* w, double* ap) {
    hipLaunchKernelGGL(kCprPressure<NB>, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, val, w, ap);
}
template <int NB>
void cprRestrictT(hipStream_t s, int n, const double* w, const double* r, double* rp, const void* dinv, bool f32, double omega, double* x1) {
    if (f32) hipLaunchKernelGGL((kCprRestrict<NB, float>), dim3((n + 255) / 256), dim3(256), 0, s, n, w, r, rp, (const float*)dinv, omega, x1);
    else hipLaunchKernelGGL((kCprRestrict<NB, double>), dim3((n + 255) / 256), dim3(256), 0, s, n, w, r, rp, (const double*)dinv, omega, x1);
}
template <int NB, typename TA, typename CT>
void cprSecondCT(hipStream_t s, const SliceDev<double>* sl, const CT* col, const TA* colP, const double* r, const double* xp, double* t,
                 const int* iperm, int w1, double* y0) {
    const int waves = sl->nslices + sl->nlong;
    hipLaunchKernelGGL((kCprSecond<NB, TA, CT>), dim3((waves + 3) / 4), dim3(kBlock), 0, s, sl->nslices, sl->info.p, sl->row.p, col, colP,
                       sl->nlong, sl->lrow.p, sl->lptr.p, sl->lcol.p, sl->lval.p, r, xp, t, iperm, w1, y0);
}
template <int NB>
void cprSecondT(hipStream_t s, const SliceDev<double>* sl, const void* colP, bool f32, const double* r, const double* xp, double* t,
                const int* iperm, int w1, double* y0) {
    if (sl->dcol.n) {
        if (f32) cprSecondCT<NB, float, short>(s, sl, sl->dcol.p, (const float*)colP, r, xp, t, iperm, w1, y0);
        else cprSecondCT<NB, double, short>(s, sl, sl->dcol.p, (const double*)colP, r, xp, t, iperm, w1, y0);
    } else {
        if (f32) cprSecondCT<NB, float, int>(s, sl, sl->col.p, (const float*)colP, r, xp, t, iperm, w1, y0);
        else cprSecondCT<NB, double, int>(s, sl, sl->col.p, (const double*)colP, r, xp, t, iperm, w1, y0);
    }
}



template <int NB, typename TA, typename TV>
void iluFusedT(hipStream_t s, const SliceDev<TA>& d, bool fwd, int l0, int l1, const int* sStart, const int* qStart, int wbase, int wcount,
               const int* perm, const double* r, TV* y, double* z, const double* xp) {
    const SliceView<TA> A{d.info.p, d.row.p, d.orig.p, d.col.p, d.val.p, d.dval.p, d.lrow.p, d.lorig.p, d.lptr.p, d.lcol.p, d.lval.p, d.ldval.p, d.owner.p, d.lowner.p};
    if (fwd) hipLaunchKernelGGL((kIluFused<NB, TA, TV, true>), dim3(1), dim3(kFuseThreads), 0, s, l0, l1, sStart, qStart, A, wbase, wcount, perm, r, y, z, xp);
    else hipLaunchKernelGGL((kIluFused<NB, TA, TV, false>), dim3(1), dim3(kFuseThreads), 0, s, l0, l1, sStart, qStart, A, wbase, wcount, perm, r, y, z, xp);
}
template <int NB, typename TA, typename TV>
void iluTileT(hipStream_t s, const SliceDev<TA>& d, bool fwd, int u0, int nu, const int* unitStart, const int* sStart, const int* qStart,
              const int* lvPtr, const int* lvStart, int rowsCap, int levelsCap, int longMax, int longCap, const int* perm, const double* r, TV* y,
              double* z, const double* xp) {
    const SliceView<TA> A{d.info.p, d.row.p, d.orig.p, d.col.p, d.val.p, d.dval.p, d.lrow.p, d.lorig.p, d.lptr.p, d.lcol.p, d.lval.p, d.ldval.p, d.owner.p, d.lowner.p};
    const int lds = tileLds(rowsCap, levelsCap, longMax, longCap, NB, int(sizeof(TV)), int(sizeof(TA))).total;
    if (fwd) hipLaunchKernelGGL((kIluTile<NB, TA, TV, true>), dim3(nu), dim3(kTileThreads), lds, s, u0, unitStart, sStart, qStart, lvPtr, lvStart, A, rowsCap,
                                levelsCap, longMax, longCap, perm, r, y, z, xp);
    else hipLaunchKernelGGL((kIluTile<NB, TA, TV, false>), dim3(nu), dim3(kTileThreads), lds, s, u0, unitStart, sStart, qStart, lvPtr, lvStart, A, rowsCap,
                            levelsCap, longMax, longCap, perm, r, y, z, xp);
}
template <int NB>
void bilu0MappedT(hipStream_t s, const int* rows, int r0, int r1, const int* rp, const int* ci, const int* diag, double* val, int* nreg, int lean,
                  const LongRowMap& m, int lds) {
    hipLaunchKernelGGL(kBilu0LongMapped<NB>, dim3(r1 - r0), dim3(kWave), lds, s, rows, r0, rp, ci, diag, val, nreg, lean, m.mrow, m.mptr, m.map);
}
template <int NB>
void bilu0LongT(hipStream_t s, const int* rows, int r0, int r1, const int* rp, const int* ci, const int* diag, double* val, int* nreg,
                int lean) {
    hipLaunchKernelGGL(kBilu0LongRows<NB>, dim3(r1 - r0), dim3(kWave), 0, s, rows, r0, rp, ci, diag, val, nreg, lean);
}

inline void checkLaunch() {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(-2, std::string("kernel launch: ") + hipGetErrorString(e));
}

#define RSH_NB_DISPATCH(nb, CALL)                                   \
    switch (nb) {                                                   \
        case 1: { constexpr int NB = 1; CALL; } break;              \
        case 2: { constexpr int NB = 2; CALL; } break;              \
        case 3: { constexpr int NB = 3; CALL; } break;              \
        case 4: { constexpr int NB = 4; CALL; } break;              \
        default: throw Error(-6, "block size not built (1..4)");    \
    }

}  // namespace

// ================================================================ launchers
void launchPackSlices(hipStream_t s, int nb, const double* src, const int* srcIdx, long long nslots, void* dst, bool f32) {
    const long long total = nslots * nb * nb * kWave;
    if (!total) return;
    if (f32) hipLaunchKernelGGL(kPackSlices<float>, dim3(gridFor(total)), dim3(kBlock), 0, s, nb * nb, src, srcIdx, nslots, (float*)dst);
    else hipLaunchKernelGGL(kPackSlices<double>, dim3(gridFor(total)), dim3(kBlock), 0, s, nb * nb, src, srcIdx, nslots, (double*)dst);
    checkLaunch();
}
void launchPackLong(hipStream_t s, int nb, const double* src, const int* srcIdx, int count, void* dst, bool f32) {
    const long long total = (long long)count * nb * nb;
    if (!total) return;
    if (f32) hipLaunchKernelGGL(kPackLong<float>, dim3(gridFor(total)), dim3(kBlock), 0, s, nb * nb, src, srcIdx, count, (float*)dst);
    else hipLaunchKernelGGL(kPackLong<double>, dim3(gridFor(total)), dim3(kBlock), 0, s, nb * nb, src, srcIdx, count, (double*)dst);
    checkLaunch();
}
void launchPackColumn0(hipStream_t s, int nb, const double* src, const int* srcIdx, long long nslots, double* dst) {
    const long long total = nslots * nb * kWave;
    if (!total) return;
    hipLaunchKernelGGL(kPackColumn0, dim3(gridFor(total)), dim3(kBlock), 0, s, nb, src, srcIdx, nslots, dst);
    checkLaunch();
}
void launchGatherBlocks(hipStream_t s, int nb, const double* src, const int* map, int count, double* dst) {
    const long long total = (long long)count * nb * nb;
    if (!total) return;
    hipLaunchKernelGGL(kGatherBlocks, dim3(gridFor(total)), dim3(kBlock), 0, s, nb * nb, src, map, count, dst);
    checkLaunch();
}
void launchSpmvSliced(hipStream_t s, int nb, bool f32, const SliceDev<double>* d64, const SliceDev<float>* d32, const double* x, double* y,
                      const KScalars* gate) {
    RSH_NB_DISPATCH(nb, (spmvT<NB>(s, f32, d64, d32, x, y, gate)));
    checkLaunch();
}
void launchIluForward(hipStream_t s, int nb, int fp, const void* slices, int s0, int s1, int q0, int q1, const double* r, void* y,
                      const KScalars* gate, double* zFused, const double* xp) {
    if (fp == kIluFp32) RSH_NB_DISPATCH(nb, (iluFwdT<NB, float, float>(s, *static_cast<const SliceDev<float>*>(slices), s0, s1, q0, q1, r, (float*)y, gate, zFused, xp)))
    else if (fp == kIluStore32) RSH_NB_DISPATCH(nb, (iluFwdT<NB, float, double>(s, *static_cast<const SliceDev<float>*>(slices), s0, s1, q0, q1, r, (double*)y, gate, zFused, xp)))
    else RSH_NB_DISPATCH(nb, (iluFwdT<NB, double, double>(s, *static_cast<const SliceDev<double>*>(slices), s0, s1, q0, q1, r, (double*)y, gate, zFused, xp)))
    checkLaunch();
}
void launchIluBackward(hipStream_t s, int nb, int fp, const void* slices, int s0, int s1, int q0, int q1, void* y, double* z, const double* xp,
                       const KScalars* gate, bool writeY) {
    if (fp == kIluFp32) RSH_NB_DISPATCH(nb, (iluBwdT<NB, float, float>(s, *static_cast<const SliceDev<float>*>(slices), s0, s1, q0, q1, (float*)y, z, xp, gate, writeY)))
    else if (fp == kIluStore32) RSH_NB_DISPATCH(nb, (iluBwdT<NB, float, double>(s, *static_cast<const SliceDev<float>*>(slices), s0, s1, q0, q1, (double*)y, z, xp, gate, writeY)))
    else RSH_NB_DISPATCH(nb, (iluBwdT<NB, double, double>(s, *static_cast<const SliceDev<double>*>(slices), s0, s1, q0, q1, (double*)y, z, xp, gate, writeY)))
    checkLaunch();
}


template <int NB, typename TA, typename TV>
void iluTileWaveT(hipStream_t s, const SliceDev<TA>& d, bool fwd, int ntiles, const int4* tiles, const double* r, TV* y, double* z, const double* xp) {
    if (ntiles <= 0) return;
    if (!fwd && !d.dval.n) throw Error(-1, "tile sweep: the upper slices carry no pivot blocks");
    const dim3 grid((ntiles + kTileWaveTiles - 1) / kTileWaveTiles);
    if (fwd)
        hipLaunchKernelGGL((kIluTileWave<NB, TA, TV, true>), grid, dim3(kTileWaveTiles * kWave), 0, s, ntiles, tiles, d.info.p, d.row.p, d.orig.p,
                           d.col.p, d.val.p, d.dval.p, r, y, z, xp);
    else
        hipLaunchKernelGGL((kIluTileWave<NB, TA, TV, false>), grid, dim3(kTileWaveTiles * kWave), 0, s, ntiles, tiles, d.info.p, d.row.p, d.orig.p,
                           d.col.p, d.val.p, d.dval.p, r, y, z, xp);
}
void launchIluTileWave(hipStream_t s, int nb, int fp, const void* slices, bool fwd, int ntiles, const int4* tiles, const double* r, void* y,
                       double* z, const double* xp) {
    if (fp == kIluFp32) RSH_NB_DISPATCH(nb, (iluTileWaveT<NB, float, float>(s, *static_cast<const SliceDev<float>*>(slices), fwd, ntiles, tiles, r, (float*)y, z, xp)))
    else if (fp == kIluStore32) RSH_NB_DISPATCH(nb, (iluTileWaveT<NB, float, double>(s, *static_cast<const SliceDev<float>*>(slices), fwd, ntiles, tiles, r, (double*)y, z, xp)))
    else RSH_NB_DISPATCH(nb, (iluTileWaveT<NB, double, double>(s, *static_cast<const SliceDev<double>*>(slices), fwd, ntiles, tiles, r, (double*)y, z, xp)))
    checkLaunch();
}

template <int NB, typename TA, typename TV>
void runOutsideT(hipStream_t s, const SliceDev<TA>& d, bool fwd, int t0, int ns, int u0, int nq, int wbase, int wend, const double* r,
                 const TV* y, TV* rt, const KScalars* gate) {
    const SliceView<TA> A{d.info.p, d.row.p, d.orig.p, d.col.p, d.val.p, d.dval.p, d.lrow.p, d.lorig.p, d.lptr.p, d.lcol.p, d.lval.p, d.ldval.p, d.owner.p, d.lowner.p};
    const int blocks = (ns + nq + 3) / 4;
    if (fwd) hipLaunchKernelGGL((kRunOutside<NB, TA, TV, true>), dim3(blocks), dim3(256), 0, s, t0, ns, u0, nq, A, wbase, wend, r, y, rt, gate);
    else hipLaunchKernelGGL((kRunOutside<NB, TA, TV, false>), dim3(blocks), dim3(256), 0, s, t0, ns, u0, nq, A, wbase, wend, r, y, rt, gate);
}
template <int NB, typename TM, typename TV>
void runDenseT(hipStream_t s, bool fwd, int ntasks, const int2* tasks, const int* start, const int* rows, const long long* off, const TM* inv,
               int wbase, const TV* rt, TV* y, double* z, const int* perm, const double* xp, const int* gowner, const KScalars* gate,
               TV* ycopy) {
    if (fwd) hipLaunchKernelGGL((kRunDense<NB, TM, TV, true>), dim3(ntasks), dim3(kDenseThreads), 0, s, tasks, start, rows, off, inv, wbase, rt, y, z, perm, xp,
                                gowner, gate, ycopy);
    else hipLaunchKernelGGL((kRunDense<NB, TM, TV, false>), dim3(ntasks), dim3(kDenseThreads), 0, s, tasks, start, rows, off, inv, wbase, rt, y, z, perm, xp,
                            gowner, gate, static_cast<TV*>(nullptr));
}
void launchRunDense(hipStream_t s, int nb, int fp, const void* slices, bool fwd, int t0, int ns, int u0, int nq, int wbase, int wend,
                    int ntasks, const int2* tasks, const int* start, const int* rows, const long long* off, const void* inv,
                    const int* perm, const double* r, void* rt, void* y, double* z, const double* xp, const int* gowner, const KScalars* gate,
                    void* ycopy, bool inv64) {
    if (ns + nq > 0) {
        if (fp == kIluFp32) RSH_NB_DISPATCH(nb, (runOutsideT<NB, float, float>(s, *static_cast<const SliceDev<float>*>(slices), fwd, t0, ns, u0, nq, wbase, wend, r, (const float*)y, (float*)rt, gate)))
        else if (fp == kIluStore32) RSH_NB_DISPATCH(nb, (runOutsideT<NB, float, double>(s, *static_cast<const SliceDev<float>*>(slices), fwd, t0, ns, u0, nq, wbase, wend, r, (const double*)y, (double*)rt, gate)))
        else RSH_NB_DISPATCH(nb, (runOutsideT<NB, double, double>(s, *static_cast<const SliceDev<double>*>(slices), fwd, t0, ns, u0, nq, wbase, wend, r, (const double*)y, (double*)rt, gate)))
        checkLaunch();
    }
    if (fp == kIluFp32) RSH_NB_DISPATCH(nb, (runDenseT<NB, float, float>(s, fwd, ntasks, tasks, start, rows, off, (const float*)inv, wbase, (const float*)rt, (float*)y, z, perm, xp, gowner, gate, (float*)ycopy)))
    else if (fp == kIluStore32 && !inv64) RSH_NB_DISPATCH(nb, (runDenseT<NB, float, double>(s, fwd, ntasks, tasks, start, rows, off, (const float*)inv, wbase, (const double*)rt, (double*)y, z, perm, xp, gowner, gate, (double*)ycopy)))
    else RSH_NB_DISPATCH(nb, (runDenseT<NB, double, double>(s, fwd, ntasks, tasks, start, rows, off, (const double*)inv, wbase, (const double*)rt, (double*)y, z, perm, xp, gowner, gate, (double*)ycopy)))
    checkLaunch();
}
// block columns per workgroup: enough workgroups per group that a step's updates are a few blocks per thread
template <int NB>
void denseInvertT(hipStream_t s, bool upper, int ngroups, int maxRows, const int* start, const long long* off, double* raw, double* inv) {
    const int P = std::max(1, std::min(maxRows, 32));
    if (upper) {
        const unsigned bx = unsigned(((long long)maxRows * maxRows + 255) / 256);
        hipLaunchKernelGGL((kDenseScaleU<NB>), dim3(std::max(1u, bx), unsigned(ngroups)), dim3(256), 0, s, start, off, raw);
        hipLaunchKernelGGL((kDenseInvert<NB, true>), dim3(unsigned(P), unsigned(ngroups)), dim3(256), 0, s, start, off, raw, inv);
    } else {
        hipLaunchKernelGGL((kDenseInvert<NB, false>), dim3(unsigned(P), unsigned(ngroups)), dim3(256), 0, s, start, off, raw, inv);
    }
}
void launchDenseGroups(hipStream_t s, int nb, bool upper, int ngroups, int maxRows, const int* start, const long long* off, int nmap,
                       const int4* map, const double* lu, double* raw, long long rawElems, double* inv) {
    if (ngroups == 0) return;
    if (ngroups > 65535) throw Error(-1, "too many dense groups for one launch");
    launchFill(s, raw, 0.0, rawElems);
    if (nmap > 0) {
        const long long t = (long long)nmap * nb * nb;
        hipLaunchKernelGGL(kDenseGather, dim3(unsigned((t + 255) / 256)), dim3(256), 0, s, nmap, nb, map, start, off, lu, raw);
        checkLaunch();
    }
    RSH_NB_DISPATCH(nb, (denseInvertT<NB>(s, upper, ngroups, maxRows, start, off, raw, inv)))
    checkLaunch();
}
void launchPerturbRows(hipStream_t s, int n, const int* rp, int nb, const double* base, double* out, double amp, unsigned long long seed) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kPerturbRows, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, rp, nb * nb, base, out, amp, seed);
    checkLaunch();
}
template <int NB>
void blockDiagInvT(hipStream_t s, int n, const int* diag, const double* val, double* out) {
    hipLaunchKernelGGL(kBlockDiagInv<NB>, dim3((n + 255) / 256), dim3(256), 0, s, n, diag, val, out);
}
void launchBlockDiagInv(hipStream_t s, int nb, int n, const int* diag, const double* val, double* out) {
    if (n <= 0) return;
    RSH_NB_DISPATCH(nb, (blockDiagInvT<NB>(s, n, diag, val, out)))
    checkLaunch();
}
template <int NB>
void blockJacobiT(hipStream_t s, int n, const void* dinv, bool f32, const double* t, const double* xp, double* z) {
    if (f32) hipLaunchKernelGGL((kBlockJacobi<NB, float>), dim3((n + 255) / 256), dim3(256), 0, s, n, (const float*)dinv, t, xp, z);
    else hipLaunchKernelGGL((kBlockJacobi<NB, double>), dim3((n + 255) / 256), dim3(256), 0, s, n, (const double*)dinv, t, xp, z);
}
void launchBlockJacobi(hipStream_t s, int nb, int n, const void* dinv, bool f32, const double* t, const double* xp, double* z) {
    if (n <= 0) return;
    RSH_NB_DISPATCH(nb, (blockJacobiT<NB>(s, n, dinv, f32, t, xp, z)))
    checkLaunch();
}
void launchWellDinv(hipStream_t s, int nw, const int* wdiag, const int* wcPtr, const int* wcSrc, const int* wwPtr, const int* wwSrc, const double* jval,
                    double* dinv, int* bad) {
    if (nw <= 0) return;
    hipLaunchKernelGGL(kWellDinv, dim3((nw + 63) / 64), dim3(64), 0, s, nw, wdiag, wcPtr, wcSrc, wwPtr, wwSrc, jval, dinv, bad);
    checkLaunch();
}
void launchSchurValues(hipStream_t s, long long nvals, const int* direct, const int* termPtr, const int4* terms, const double* jval,
                       const double* dinv, double* sval) {
    if (nvals <= 0) return;
    hipLaunchKernelGGL(kSchurValues, dim3(unsigned((nvals + 255) / 256)), dim3(256), 0, s, nvals, direct, termPtr, terms, jval, dinv, sval);
    checkLaunch();
}
void launchSchurRhs(hipStream_t s, int nc, const int* cwPtr, const int* cwWell, const int* cwSrc, const double* jval, const double* dinv,
                    const double* b, double* bc) {
    if (nc <= 0) return;
    hipLaunchKernelGGL(kSchurRhs, dim3((nc + 255) / 256), dim3(256), 0, s, nc, cwPtr, cwWell, cwSrc, jval, dinv, b, bc);
    checkLaunch();
}
void launchSchurRecover(hipStream_t s, int nc, int nw, const int* wcPtr, const int* wcCol, const int* wcSrc, const double* jval, const double* dinv,
                        const double* b, const double* xc, double* x) {
    const int m = std::max(nc, nw);
    if (m <= 0) return;
    hipLaunchKernelGGL(kSchurRecover, dim3((m + 255) / 256), dim3(256), 0, s, nc, nw, wcPtr, wcCol, wcSrc, jval, dinv, b, xc, x);
    checkLaunch();
}
void launchToFloat(hipStream_t s, long long n, const double* a, float* b) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kToFloat, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, n, a, b);
    checkLaunch();
}

void launchIluFused(hipStream_t s, int nb, int fp, const void* slices, bool fwd, int l0, int l1, const int* sStart, const int* qStart,
                    int wbase, int wcount, const int* perm, const double* r, void* y, double* z, const double* xp) {
    if (l1 <= l0) return;
    if (wcount > kRunRowsMax) throw Error(-1, "fused level run longer than the LDS buffer");
    if (fp == kIluFp32) RSH_NB_DISPATCH(nb, (iluFusedT<NB, float, float>(s, *static_cast<const SliceDev<float>*>(slices), fwd, l0, l1, sStart, qStart, wbase, wcount, perm, r, (float*)y, z, xp)))
    else if (fp == kIluStore32) RSH_NB_DISPATCH(nb, (iluFusedT<NB, float, double>(s, *static_cast<const SliceDev<float>*>(slices), fwd, l0, l1, sStart, qStart, wbase, wcount, perm, r, (double*)y, z, xp)))
    else RSH_NB_DISPATCH(nb, (iluFusedT<NB, double, double>(s, *static_cast<const SliceDev<double>*>(slices), fwd, l0, l1, sStart, qStart, wbase, wcount, perm, r, (double*)y, z, xp)))
    checkLaunch();
}
int tileLdsBytes(int nb, int fp, int rowsCap, int levelsCap, int longMax, int longCap) {
    const int tv = fp == kIluFp32 ? 4 : 8, ta = fp == kIluFp64 ? 8 : 4;
    return tileLds(rowsCap, levelsCap, longMax, longCap, nb, tv, ta).total;
}
void launchIluTile(hipStream_t s, int nb, int fp, const void* slices, bool fwd, int u0, int nu, const int* unitStart, const int* sStart,
                   const int* qStart, const int* lvPtr, const int* lvStart, int rowsCap, int levelsCap, int longMax, int longCap, const int* perm,
                   const double* r, void* y, double* z, const double* xp) {
    if (nu <= 0) return;
    if (tileLdsBytes(nb, fp, rowsCap, levelsCap, longMax, longCap) > kTileLdsMax) throw Error(-1, "tile sweep: LDS request above the limit");
    if (fp == kIluFp32) RSH_NB_DISPATCH(nb, (iluTileT<NB, float, float>(s, *static_cast<const SliceDev<float>*>(slices), fwd, u0, nu, unitStart, sStart, qStart, lvPtr, lvStart, rowsCap, levelsCap, longMax, longCap, perm, r, (float*)y, z, xp)))
    else if (fp == kIluStore32) RSH_NB_DISPATCH(nb, (iluTileT<NB, float, double>(s, *static_cast<const SliceDev<float>*>(slices), fwd, u0, nu, unitStart, sStart, qStart, lvPtr, lvStart, rowsCap, levelsCap, longMax, longCap, perm, r, (double*)y, z, xp)))
    else RSH_NB_DISPATCH(nb, (iluTileT<NB, double, double>(s, *static_cast<const SliceDev<double>*>(slices), fwd, u0, nu, unitStart, sStart, qStart, lvPtr, lvStart, rowsCap, levelsCap, longMax, longCap, perm, r, (double*)y, z, xp)))
    checkLaunch();
}
void launchGsFused(hipStream_t s, int l0, int l1, const int* lstart, const int* rows, const int* rp, const int* ci, const int* diag, const double* v,
                   const double* b, const double* xin, double* xout, bool forward) {
    if (l1 <= l0) return;
    hipLaunchKernelGGL(kGsFused, dim3(1), dim3(kFuseThreads), 0, s, l0, l1, lstart, rows, rp, ci, diag, v, b, xin, xout, forward ? 1 : 0);
    checkLaunch();
}
void launchBilu0LongRows(hipStream_t s, int nb, const int* rows, int r0, int r1, const int* rp, const int* ci, const int* diag, double* val, int* nreg,
                         bool lean, const LongRowMap* map) {
    if (r1 <= r0) return;
    if (map && map->maxLen > 0) {
        const int lds = map->maxLen * nb * nb * int(sizeof(double));
        if (lds > kLongRowLdsMax) throw Error(-1, "long-row elimination: row longer than the LDS buffer");
        RSH_NB_DISPATCH(nb, (bilu0MappedT<NB>(s, rows, r0, r1, rp, ci, diag, val, nreg, lean ? 1 : 0, *map, lds)));
    } else {
        RSH_NB_DISPATCH(nb, (bilu0LongT<NB>(s, rows, r0, r1, rp, ci, diag, val, nreg, lean ? 1 : 0)));
    }
    checkLaunch();
}
void launchBilu0Level(hipStream_t s, int nb, const int* rows, int r0, int r1, const int* rp, const int* ci, const int* diag, double* val, int* nreg,
                      int stop, bool lean) {
    const int cnt = r1 - r0;
    if (cnt <= 0) return;
    RSH_NB_DISPATCH(nb, (bilu0LevelT<NB>(s, rows, r0, r1, rp, ci, diag, val, nreg, stop, lean ? 1 : 0)));
    checkLaunch();
}
void launchIluGroupFactor(hipStream_t s, int nb, int g0, int g1, const int* gstart, const int* grows, const long long* gidxOff, const int* gidx,
                          const int* diag, double* val, int* nreg, const int* rp, bool lean) {
    if (g1 <= g0) return;
    RSH_NB_DISPATCH(nb, (iluGroupFactorT<NB>(s, g0, g1, gstart, grows, gidxOff, gidx, diag, val, nreg, rp, lean ? 1 : 0)));
    checkLaunch();
}
void launchRoundToFloat(hipStream_t s, long long n, double* x) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kRoundToFloat, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, n, x);
    checkLaunch();
}
void launchCprWeights(hipStream_t s, int nb, int n, const int* diag, const double* val, double* w, const double* acc, int* fallbacks) {
    RSH_NB_DISPATCH(nb, (cprWeightsT<NB>(s, n, diag, val, w, acc, fallbacks)));
    checkLaunch();
}
void launchCprPressure(hipStream_t s, int nb, int n, const int* rp, const double* val, const double* w, double* ap) {
    RSH_NB_DISPATCH(nb, (cprPressureT<NB>(s, n, rp, val, w, ap)));
    checkLaunch();
}
void launchFixDiag(hipStream_t s, int n, const int* diag, double* v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kFixDiag, dim3((n + 255) / 256), dim3(256), 0, s, n, diag, v);
    checkLaunch();
}
void launchCprRestrict(hipStream_t s, int nb, int n, const double* w, const double* r, double* rp, const void* dinv, bool f32, double omega,
                       double* x1) {
    if (!x1) dinv = nullptr;
    RSH_NB_DISPATCH(nb, (cprRestrictT<NB>(s, n, w, r, rp, dinv, f32, omega, x1)));
    checkLaunch();
}
void launchCprSecond(hipStream_t s, int nb, const SliceDev<double>* sl, const void* colP, bool f32, const double* r, const double* xp, double* t,
                     const int* iperm, int w1, double* y0) {
    const int waves = sl->nslices + sl->nlong;
    if (!waves) return;
    if (!y0) iperm = nullptr;
    RSH_NB_DISPATCH(nb, (cprSecondT<NB>(s, sl, colP, f32, r, xp, t, iperm, w1, y0)));
    checkLaunch();
}
void launchCsrSpmvResidual(hipStream_t s, int n, const int* rp, const int* ci, const double* v, const double* x, const double* b, double* r) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kCsrResidual, dim3((n + 255) / 256), dim3(256), 0, s, n, rp, ci, v, x, b, r);
    checkLaunch();
}
void launchGsLevel(hipStream_t s, int nrows, const int* rows, const int* rp, const int* ci, const int* diag, const double* v, const double* b,
                   const double* xin, double* xout, bool forward) {
    if (nrows <= 0) return;
    hipLaunchKernelGGL(kGsLevel, dim3((nrows + 127) / 128), dim3(128), 0, s, nrows, rows, rp, ci, diag, v, b, xin, xout, forward ? 1 : 0);
    checkLaunch();
}
template <int G>
void restrictT(hipStream_t s, int nc, const int* rrp, const int* rci, const double* r, double* bc, const void* dinv, bool f32, double omega,
               double* x1) {
    const long long th = (long long)nc * G;
    const dim3 grid(unsigned((th + 255) / 256));
    if (f32) hipLaunchKernelGGL((kRestrict<G, float>), grid, dim3(256), 0, s, nc, rrp, rci, r, bc, (const float*)dinv, omega, x1);
    else hipLaunchKernelGGL((kRestrict<G, double>), grid, dim3(256), 0, s, nc, rrp, rci, r, bc, (const double*)dinv, omega, x1);
}
template <int MODE, int G, typename TA>
void jacobiCsrT(hipStream_t s, int n, const int* rp, const int* ci, const TA* v, const TA* dinv, const double* b, const double* x1,
                double omega, double* out0, const int* prp, const int* pci, double* pfine) {
    const long long th = (long long)n * G;
    hipLaunchKernelGGL((kJacobiCsr<MODE, G, TA>), dim3(unsigned((th + 255) / 256)), dim3(256), 0, s, n, rp, ci, v, dinv, b, x1, omega, out0, prp,
                       pci, pfine);
}
// lanes per row for a mean row length
int groupLanes(double mean) {
    int g = 4;
    while (g < 64 && g < mean) g *= 2;
    return g;
}
#define RSH_G_DISPATCH(g, CALL)                                  \
    switch (g) {                                                 \
        case 4: { constexpr int G = 4; CALL; } break;            \
        case 8: { constexpr int G = 8; CALL; } break;            \
        case 16: { constexpr int G = 16; CALL; } break;          \
        case 32: { constexpr int G = 32; CALL; } break;          \
        default: { constexpr int G = 64; CALL; } break;          \
    }
void launchRestrict(hipStream_t s, int nc, const int* rrp, const int* rci, const double* r, double* bc, double meanAgg, const void* dinv,
                    bool f32, double omega, double* x1) {
    if (nc <= 0) return;
    if (!x1) dinv = nullptr;
    RSH_G_DISPATCH(groupLanes(meanAgg), (restrictT<G>(s, nc, rrp, rci, r, bc, dinv, f32, omega, x1)))
    checkLaunch();
}
void launchJacobiInit(hipStream_t s, int n, bool f32, const void* dinv, const double* b, double omega, double* x1) {
    if (n <= 0) return;
    if (f32) hipLaunchKernelGGL(kJacobiInit<float>, dim3((n + 255) / 256), dim3(256), 0, s, n, (const float*)dinv, b, omega, x1);
    else hipLaunchKernelGGL(kJacobiInit<double>, dim3((n + 255) / 256), dim3(256), 0, s, n, (const double*)dinv, b, omega, x1);
    checkLaunch();
}
void launchProlongAdd(hipStream_t s, int n, const int* agg, const double* xc, double* x) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kProlongAdd, dim3((n + 255) / 256), dim3(256), 0, s, n, agg, xc, x);
    checkLaunch();
}
void launchDenseGemv(hipStream_t s, int m, const double* M, const double* x, double* y) { launchDenseGemv2(s, m, m, M, x, y); }
void launchDenseGemv2(hipStream_t s, int m, int ld, const double* M, const double* x, double* y, const int* prp, const int* pci, double* pfine) {
    if (m <= 0) return;
    hipLaunchKernelGGL(kDenseGemv, dim3((m + 3) / 4), dim3(kBlock), 0, s, m, ld, M, x, y, prp, pci, pfine);
    checkLaunch();
}
void launchGalerkin(hipStream_t s, int nap, const int* apPtr, const int* apSrc, const double* fine, double* ap, int nnzc, const int* cPtr,
                    const int* cSrc, double* coarse) {
    if (nap > 0) hipLaunchKernelGGL(kGalerkinSum, dim3((nap + 255) / 256), dim3(256), 0, s, nap, apPtr, apSrc, fine, ap);
    if (nnzc > 0) hipLaunchKernelGGL(kGalerkinSum, dim3((nnzc + 255) / 256), dim3(256), 0, s, nnzc, cPtr, cSrc, ap, coarse);
    checkLaunch();
}
void launchDenseFromCsr(hipStream_t s, int m, const int* rp, const int* ci, const double* v, double* M) {
    const long long total = (long long)m * 2 * m;
    if (!total) return;
    hipLaunchKernelGGL(kDenseFromCsr, dim3(gridFor(total)), dim3(kBlock), 0, s, m, rp, ci, v, M);
    checkLaunch();
}
void launchGaussJordanStep(hipStream_t s, int m, int k, int steps, const double* Min, double* Mout) {
    const long long total = (long long)m * 2 * m;
    int g = int(std::min<long long>((total + kBlock * 4 - 1) / (kBlock * 4), 2048));
    if (steps == 2) {
        if (k + 1 >= m) throw Error(-1, "two Gauss-Jordan steps past the last column");
        hipLaunchKernelGGL(kGaussJordanStep2, dim3(g), dim3(kBlock), 0, s, m, k, Min, Mout);
    } else {
        hipLaunchKernelGGL(kGaussJordanStep, dim3(g), dim3(kBlock), 0, s, m, k, Min, Mout);
    }
    checkLaunch();
}

template <int MODE, typename TA>
void jacobiSlicedT(hipStream_t s, const SliceDev<double>* sl, const TA* sval, const TA* lval, const TA* dinv, const double* b,
                   const double* x1, double omega, double* out0) {
    const int waves = sl->nslices + sl->nlong;
    if (!waves) return;
    const dim3 grid((waves + 3) / 4);
    if (sl->dcol.n)
        hipLaunchKernelGGL((kJacobiSliced<MODE, TA, short>), grid, dim3(kBlock), 0, s, sl->nslices, sl->info.p, sl->row.p, sl->dcol.p, sval,
                           sl->nlong, sl->lrow.p, sl->lptr.p, sl->lcol.p, lval, dinv, b, x1, omega, out0);
    else
        hipLaunchKernelGGL((kJacobiSliced<MODE, TA, int>), grid, dim3(kBlock), 0, s, sl->nslices, sl->info.p, sl->row.p, sl->col.p, sval,
                           sl->nlong, sl->lrow.p, sl->lptr.p, sl->lcol.p, lval, dinv, b, x1, omega, out0);
}
template <int MODE, int G, typename TA>
void jacobiSlicedGT(hipStream_t s, const SliceDev<double>* sl, const TA* sval, const TA* lval, const TA* dinv, const double* b,
                    const double* x1, double omega, double* out0) {
    const int waves = sl->nslices + (sl->nlong + kWave / G - 1) / (kWave / G);
    if (!waves) return;
    const dim3 grid((waves + 3) / 4);
    if (sl->dcol.n)
        hipLaunchKernelGGL((kJacobiSlicedG<MODE, G, TA, short>), grid, dim3(kBlock), 0, s, sl->nslices, sl->info.p, sl->row.p, sl->dcol.p, sval,
                           sl->nlong, sl->lrow.p, sl->lptr.p, sl->lcol.p, lval, dinv, b, x1, omega, out0);
    else
        hipLaunchKernelGGL((kJacobiSlicedG<MODE, G, TA, int>), grid, dim3(kBlock), 0, s, sl->nslices, sl->info.p, sl->row.p, sl->col.p, sval,
                           sl->nlong, sl->lrow.p, sl->lptr.p, sl->lcol.p, lval, dinv, b, x1, omega, out0);
}
template <typename TA>
void jacobiFusedT(hipStream_t s, int mode, const SliceDev<double>* sl, const TA* sval, const TA* lval, int n, const int* rp, const int* ci,
                  const TA* v, const TA* dinv, const double* b, const double* x1, double omega, double* out0, double meanRow,
                  const int* prp, const int* pci, double* pfine) {
    if (sl) {
        if (pfine) throw Error(-1, "a sliced Jacobi step does not prolong: its caller adds P x");
        if (sl->chains) {
            if (sl->chains != 8 && sl->chains != 16 && sl->chains != 32) throw Error(-1, "sliced coarse level with an unsupported lane count");
            if (mode == 0) {
                if (sl->chains == 8) jacobiSlicedGT<0, 8, TA>(s, sl, sval, lval, dinv, b, x1, omega, out0);
                else if (sl->chains == 16) jacobiSlicedGT<0, 16, TA>(s, sl, sval, lval, dinv, b, x1, omega, out0);
                else jacobiSlicedGT<0, 32, TA>(s, sl, sval, lval, dinv, b, x1, omega, out0);
            } else {
                if (sl->chains == 8) jacobiSlicedGT<1, 8, TA>(s, sl, sval, lval, dinv, b, x1, omega, out0);
                else if (sl->chains == 16) jacobiSlicedGT<1, 16, TA>(s, sl, sval, lval, dinv, b, x1, omega, out0);
                else jacobiSlicedGT<1, 32, TA>(s, sl, sval, lval, dinv, b, x1, omega, out0);
            }
            return;
        }
        if (mode == 0) jacobiSlicedT<0, TA>(s, sl, sval, lval, dinv, b, x1, omega, out0);
        else jacobiSlicedT<1, TA>(s, sl, sval, lval, dinv, b, x1, omega, out0);
        return;
    }
    if (n <= 0) return;
    const int g = groupLanes(meanRow);
    if (mode == 0) RSH_G_DISPATCH(g, (jacobiCsrT<0, G, TA>(s, n, rp, ci, v, dinv, b, x1, omega, out0, prp, pci, pfine)))
    else RSH_G_DISPATCH(g, (jacobiCsrT<1, G, TA>(s, n, rp, ci, v, dinv, b, x1, omega, out0, prp, pci, pfine)))
}
void launchJacobiFused(hipStream_t s, int mode, bool f32, const SliceDev<double>* sl, const void* sval, const void* lval, int n, const int* rp,
                       const int* ci, const void* v, const void* dinv, const double* b, const double* x1, double omega, double* out0,
                       double meanRow, const int* prp, const int* pci, double* pfine) {
    if (mode < 0 || mode > 1) throw Error(-1, "unknown Jacobi step");
    if (f32)
        jacobiFusedT<float>(s, mode, sl, (const float*)sval, (const float*)lval, n, rp, ci, (const float*)v, (const float*)dinv, b, x1, omega, out0,
                            meanRow, prp, pci, pfine);
    else
        jacobiFusedT<double>(s, mode, sl, (const double*)sval, (const double*)lval, n, rp, ci, (const double*)v, (const double*)dinv, b, x1, omega,
                             out0, meanRow, prp, pci, pfine);
    checkLaunch();
}
void launchDinv(hipStream_t s, int n, const int* diag, const double* v, double* dinv) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kDinv, dim3((n + 255) / 256), dim3(256), 0, s, n, diag, v, dinv);
    checkLaunch();
}
void launchFill(hipStream_t s, double* x, double v, long long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kFill, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, x, v, n);
    checkLaunch();
}
void launchCopy(hipStream_t s, double* dst, const double* src, long long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(kCopy, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, dst, src, n);
    checkLaunch();
}
static int redGrid(long long n) {
    long long g = (n + kBlock * 8 - 1) / (kBlock * 8);
    if (g < 1) g = 1;
    if (g > kRedBlocks) g = kRedBlocks;
    return int(g);
}
void launchDots(hipStream_t s, double* red, long long n, int k, const double* a0, const double* b0, const double* a1, const double* b1, double* out,
                int ep, KScalars* ks) {
    const int g = redGrid(n);
    if (k == 1) hipLaunchKernelGGL(kDots<1>, dim3(g), dim3(kBlock), 0, s, n, a0, b0, a1, b1, red);
    else hipLaunchKernelGGL(kDots<2>, dim3(g), dim3(kBlock), 0, s, n, a0, b0, a1, b1, red);
    hipLaunchKernelGGL(kReduceFinal, dim3(1), dim3(kBlock), 0, s, k, g, red, out, ep, ks);
    checkLaunch();
}
void launchDots(hipStream_t s, Ctx& c, long long n, int k, const double* a0, const double* b0, const double* a1, const double* b1, double* out,
                int ep, KScalars* ks) {
    launchDots(s, c.red.p, n, k, a0, b0, a1, b1, out, ep, ks);
}
void launchBicgP(hipStream_t s, long long n, const KScalars* ks, const double* r, const double* v, double* p) {
    hipLaunchKernelGGL(kBicgP, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, n, ks, r, v, p);
    checkLaunch();
}
void launchBicgS(hipStream_t s, double* red, long long n, KScalars* ks, const double* r, const double* v, double* sv) {
    const int g = redGrid(n);
    hipLaunchKernelGGL(kBicgS, dim3(g), dim3(kBlock), 0, s, n, ks, r, v, sv, red);
    hipLaunchKernelGGL(kReduceFinal, dim3(1), dim3(kBlock), 0, s, 1, g, red, nullptr, int(kEpSn), ks);
    checkLaunch();
}
void launchBicgS(hipStream_t s, Ctx& c, long long n, KScalars* ks, const double* r, const double* v, double* sv) {
    launchBicgS(s, c.red.p, n, ks, r, v, sv);
}
void launchBicgXR(hipStream_t s, double* red, long long n, KScalars* ks, const double* ph, const double* sh, const double* sv, const double* t,
                  const double* rh, double* x, double* r) {
    const int g = redGrid(n);
    hipLaunchKernelGGL(kBicgXR, dim3(g), dim3(kBlock), 0, s, n, ks, ph, sh, sv, t, rh, x, r, red);
    hipLaunchKernelGGL(kReduceFinal, dim3(1), dim3(kBlock), 0, s, 2, g, red, nullptr, int(kEpRnRho), ks);
    checkLaunch();
}
void launchBicgXR(hipStream_t s, Ctx& c, long long n, KScalars* ks, const double* ph, const double* sh, const double* sv, const double* t,
                  const double* rh, double* x, double* r) {
    launchBicgXR(s, c.red.p, n, ks, ph, sh, sv, t, rh, x, r);
}
void launchAxpy(hipStream_t s, long long n, double a, const double* x, double* y) {
    hipLaunchKernelGGL(kAxpy, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, n, a, x, y);
    checkLaunch();
}
void launchAxpyDev(hipStream_t s, long long n, const double* a, double sign, const double* x, double* y) {
    hipLaunchKernelGGL(kAxpyDev, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, n, a, sign, x, y);
    checkLaunch();
}
void launchScaleTo(hipStream_t s, long long n, const double* num, const double* x, double* y, bool invert) {
    hipLaunchKernelGGL(kScaleTo, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, n, num, x, y, invert ? 1 : 0);
    checkLaunch();
}
void launchSub(hipStream_t s, long long n, const double* b, const double* ax, double* r) {
    hipLaunchKernelGGL(kSub, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, n, b, ax, r);
    checkLaunch();
}
// k > 32 vectors: consecutive chunks of 32 (the reduction buffer holds 32 partial rows)
void launchMultiDot(hipStream_t s, Ctx& c, long long n, int k, const double* V, long long ld, const double* w, double* out) {
    if (k < 1) throw Error(-1, "multi-dot width out of range");
    const int g = redGrid(n);
    for (int q0 = 0; q0 < k; q0 += 32) {
        const int kc = std::min(32, k - q0);
        hipLaunchKernelGGL(kMultiDot, dim3(g), dim3(kBlock), 0, s, n, kc, V + (long long)q0 * ld, ld, w, c.red.p);
        hipLaunchKernelGGL(kReduceFinal, dim3(1), dim3(kBlock), 0, s, kc, g, c.red.p, out + q0, int(kEpNone), nullptr);
    }
    checkLaunch();
}
void launchMultiAxpy(hipStream_t s, long long n, int k, const double* V, long long ld, const double* h, double* w) {
    for (int q0 = 0; q0 < k; q0 += 32)
        hipLaunchKernelGGL(kMultiAxpy, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, n, std::min(32, k - q0), V + (long long)q0 * ld, ld, h + q0, w, -1.0);
    checkLaunch();
}
void launchMultiAxpyAdd(hipStream_t s, long long n, int k, const double* Z, long long ld, const double* y, double* x) {
    for (int q0 = 0; q0 < k; q0 += 32)
        hipLaunchKernelGGL(kMultiAxpy, dim3(gridFor(n, kBlock * 4)), dim3(kBlock), 0, s, n, std::min(32, k - q0), Z + (long long)q0 * ld, ld, y + q0, x, 1.0);
    checkLaunch();
}
void launchCopyKernel(hipStream_t s, const double* src, double* dst, long long n) {
    const long long n2 = n / 2;
    hipLaunchKernelGGL(kCopyVec, dim3(gridFor(n2, kBlock * 4)), dim3(kBlock), 0, s, (const double2*)src, (double2*)dst, n2);
    checkLaunch();
}

// ---------------------------------------------------------------- batched BiCGSTAB launchers
void launchBInit(hipStream_t s, int ns, const double* tol, KScalars* ks) {
    hipLaunchKernelGGL(kBInit, dim3((ns + 255) / 256), dim3(256), 0, s, ns, tol, ks);
    checkLaunch();
}
void launchBFinal(hipStream_t s, int ns, int K, const int* tStart, const double* partial, int T, int ep, KScalars* ks, double* out) {
    hipLaunchKernelGGL(kBFinal, dim3((ns + 255) / 256), dim3(256), 0, s, ns, K, tStart, partial, T, ep, ks, out);
    checkLaunch();
}
void launchBMask(hipStream_t s, int T, const int4* task, const int* keep, double* v) {
    hipLaunchKernelGGL(kBMask, dim3(T), dim3(kBlock), 0, s, task, keep, v);
    checkLaunch();
}
void launchBAdd(hipStream_t s, int T, const int4* task, const double* d, double* x) {
    hipLaunchKernelGGL(kBAdd, dim3(T), dim3(kBlock), 0, s, task, d, x);
    checkLaunch();
}
void launchBDots(hipStream_t s, int T, const int4* task, int K, const double* a0, const double* b0, const double* a1, const double* b1,
                 double* partial) {
    if (K == 1) hipLaunchKernelGGL(kBDots<1>, dim3(T), dim3(kBlock), 0, s, task, T, a0, b0, a1, b1, partial);
    else hipLaunchKernelGGL(kBDots<2>, dim3(T), dim3(kBlock), 0, s, task, T, a0, b0, a1, b1, partial);
    checkLaunch();
}
void launchBBicgP(hipStream_t s, int T, const int4* task, const KScalars* ks, const double* r, const double* v, double* p) {
    hipLaunchKernelGGL(kBBicgP, dim3(T), dim3(kBlock), 0, s, task, ks, r, v, p);
    checkLaunch();
}
void launchBBicgS(hipStream_t s, int T, const int4* task, const KScalars* ks, const double* r, const double* v, double* sv, double* partial) {
    hipLaunchKernelGGL(kBBicgS, dim3(T), dim3(kBlock), 0, s, task, T, ks, r, v, sv, partial);
    checkLaunch();
}
void launchBBicgXs(hipStream_t s, int T, const int4* task, const KScalars* ks, const double* ph, double* x) {
    hipLaunchKernelGGL(kBBicgXs, dim3(T), dim3(kBlock), 0, s, task, ks, ph, x);
    checkLaunch();
}
void launchBBicgXR(hipStream_t s, int T, const int4* task, const KScalars* ks, const double* ph, const double* sh, const double* sv,
                   const double* t, double* x, double* r, double* partial) {
    hipLaunchKernelGGL(kBBicgXR, dim3(T), dim3(kBlock), 0, s, task, T, ks, ph, sh, sv, t, x, r, partial);
    checkLaunch();
}

}  // namespace rsh
