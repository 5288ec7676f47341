// Device-side packer: the reference's MicroKernelPacker / SOPTile::pack_panel
// moved onto the GPU.  CSR input is packed in memory proportional to its
// nonzeros (no dense staging):
//
//  cell_keys       one key per CSR entry: its cell (panel, column) as
//                  panel * k + column, and its row inside the panel
//  (radix sort)    entries sorted by cell, stably: a cell's entries stay in
//                  row order
//  cell_heads      first entry of every cell; a second entry of one (row,
//                  column) is flagged as a duplicate
//  cell_fill       per cell: key, first entry, nonzero code (bit r = row r of
//                  the panel, sbench/SOP/sop_utils.py:25-30)
//  segment_bounds  first cell of every (panel, K tile) segment
//  count_cells     per segment: split every code into kernel ids through the
//                  mapping LUT (sbench/SOP/sop_driver.py:58-65) and count
//                  packed columns per id (nkern_counts, template:163-164)
//  exclusive_scan  per-segment packed column / value offsets
//  fill_cells      stable placement by id (the std::stable_sort of
//                  template:63-73) and value compression (template:158-205)
//
// Stability: columns are visited in chunks of 64 in ascending order.  Inside
// a chunk the wave repeatedly takes the first lane that still has an entry
// left, and places, for that entry's id, every lane's matching entry at
// cursor[id] + (rank among those lanes).  Each (chunk, id) pair is placed in
// one step, lanes in column order, so within an id the packed order is
// ascending column order - the same permutation a stable sort by id gives.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "srt_internal.h"

namespace srt {

namespace {

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

__device__ __forceinline__ void block_add_i64(unsigned long long *dst, int64_t v) {
    atomicAdd(dst, (unsigned long long)v);
}

// One workgroup per segment (panel, K tile): LDS histogram over kernel ids
// of the segment's cells.
__global__ void count_cells_kernel(int NK, const int32_t *__restrict__ lut_ptr,
                                   const int32_t *__restrict__ lut_ids,
                                   const int32_t *__restrict__ id_pattern,
                                   const int64_t *__restrict__ seg_cells,
                                   const uint32_t *__restrict__ cell_code,
                                   int32_t *__restrict__ counts, int64_t *__restrict__ ncols,
                                   int64_t *__restrict__ nvals) {
    extern __shared__ int hist[];
    __shared__ unsigned long long tot_c, tot_v;
    const int q = blockIdx.x;
    for (int i = threadIdx.x; i < NK; i += blockDim.x) hist[i] = 0;
    if (threadIdx.x == 0) tot_c = tot_v = 0;
    __syncthreads();
    const int64_t c0 = seg_cells[q], c1 = seg_cells[q + 1];
    for (int64_t c = c0 + threadIdx.x; c < c1; c += blockDim.x) {
        const uint32_t code = cell_code[c];
        for (int e = lut_ptr[code]; e < lut_ptr[code + 1]; ++e) atomicAdd(&hist[lut_ids[e]], 1);
    }
    __syncthreads();
    int64_t lc = 0, lv = 0;
    for (int i = threadIdx.x; i < NK; i += blockDim.x) {
        const int h = hist[i];
        counts[(int64_t)q * NK + i] = h;
        lc += h;
        lv += (int64_t)h * __popc((unsigned)id_pattern[i]);
    }
    block_add_i64(&tot_c, lc);
    block_add_i64(&tot_v, lv);
    __syncthreads();
    if (threadIdx.x == 0) {
        ncols[q] = (int64_t)tot_c;
        nvals[q] = (int64_t)tot_v;
    }
}

// Single-workgroup exclusive scan: out[0..n] with out[n] = total.
__global__ void exclusive_scan_kernel(int n, const int64_t *__restrict__ in,
                                      int64_t *__restrict__ out) {
    __shared__ int64_t sums[1024];
    const int t = threadIdx.x, T = blockDim.x;
    const int chunk = (n + T - 1) / T;
    const int beg = min(n, t * chunk), end = min(n, beg + chunk);
    int64_t s = 0;
    for (int i = beg; i < end; ++i) s += in[i];
    sums[t] = s;
    __syncthreads();
    for (int off = 1; off < T; off <<= 1) {
        const int64_t v = t >= off ? sums[t - off] : 0;
        __syncthreads();
        sums[t] += v;
        __syncthreads();
    }
    int64_t run = t ? sums[t - 1] : 0;
    for (int i = beg; i < end; ++i) {
        out[i] = run;
        run += in[i];
    }
    if (t == T - 1) out[n] = sums[T - 1];
}

// Per-id start cursors (exclusive prefix over ids) for one panel, one wave.
__device__ void id_cursors(int NK, const int32_t *__restrict__ cnt,
                           const int32_t *__restrict__ id_pattern, int *cur_c, int *cur_v) {
    const int lane = threadIdx.x & 63;
    int carry_c = 0, carry_v = 0;
    for (int base = 0; base < NK; base += 64) {
        const int i = base + lane;
        const int c = i < NK ? cnt[i] : 0;
        const int v = i < NK ? c * __popc((unsigned)id_pattern[i]) : 0;
        int sc = c, sv = v;
        for (int off = 1; off < 64; off <<= 1) {
            const int tc = __shfl_up(sc, off, 64);
            const int tv = __shfl_up(sv, off, 64);
            if (lane >= off) {
                sc += tc;
                sv += tv;
            }
        }
        if (i < NK) {
            cur_c[i] = carry_c + sc - c;
            cur_v[i] = carry_v + sv - v;
        }
        carry_c += __shfl(sc, 63, 64);
        carry_v += __shfl(sv, 63, 64);
    }
}

// Place one chunk of up to 64 lanes, each with a bitmask of pending entries
// (entry j has id ids_of(j)); emit(lane_entry j, id, rank) for every entry.
// Cursors live in LDS; one wave per workgroup.
template <typename IdOf, typename Emit>
__device__ void place_chunk(uint32_t pending, IdOf ids_of, int *cur_c, int *cur_v,
                            const int32_t *__restrict__ id_pattern, Emit emit) {
    const int lane = threadIdx.x & 63;
    while (true) {
        const uint64_t active = __ballot(pending != 0);
        if (!active) break;
        const int leader = __ffsll((long long)active) - 1;
        const int mine = pending ? ids_of(__ffs(pending) - 1) : 0;
        const int lid = __shfl(mine, leader, 64);
        int jm = -1;
        for (uint32_t q = pending; q; q &= q - 1) {
            const int j = __ffs(q) - 1;
            if (ids_of(j) == lid) {
                jm = j;
                break;
            }
        }
        const uint64_t m = __ballot(jm >= 0);
        const int rank = __popcll(m & lanemask_lt());
        const int base_c = cur_c[lid];
        const int base_v = cur_v[lid];
        const int pc = __popc((unsigned)id_pattern[lid]);
        if (jm >= 0) {
            pending &= ~(1u << jm);
            emit(jm, lid, base_c + rank, base_v + rank * pc);
        }
        __syncthreads();
        if (lane == 0) {
            const int nm = __popcll(m);
            cur_c[lid] = base_c + nm;
            cur_v[lid] = base_v + nm * pc;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int64_t wave_sum(int64_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// One 64-lane workgroup per segment (panel p = q / KT): its cells in column
// order, 64 at a time.  A cell's entries are in row order, so the value of
// row r of a kernel pattern is the next entry if that entry is row r, else a
// stored zero (padding).
__global__ void __launch_bounds__(64)
    fill_cells_kernel(int64_t k, int KT, int NK, int M_r, const int32_t *__restrict__ lut_ptr,
                      const int32_t *__restrict__ lut_ids, const int32_t *__restrict__ id_pattern,
                      const int64_t *__restrict__ seg_cells, const uint64_t *__restrict__ cell_key,
                      const uint32_t *__restrict__ cell_code, const int64_t *__restrict__ cell_first,
                      const uint32_t *__restrict__ ent_idx, const uint8_t *__restrict__ ent_row,
                      const float *__restrict__ Lx, const int64_t *__restrict__ col_ptr,
                      const int64_t *__restrict__ val_ptr, const int32_t *__restrict__ counts,
                      int32_t *__restrict__ cols, float *__restrict__ vals, int64_t *__restrict__ padding) {
    extern __shared__ int cur[];
    int *cur_c = cur, *cur_v = cur + NK;
    const int q = blockIdx.x;
    const int lane = threadIdx.x & 63;
    id_cursors(NK, counts + (int64_t)q * NK, id_pattern, cur_c, cur_v);
    __syncthreads();
    const int64_t cbase = col_ptr[q], vbase = val_ptr[q];
    const uint64_t panel_base = (uint64_t)(q / KT) * (uint64_t)k;
    const int64_t s0 = seg_cells[q], s1 = seg_cells[q + 1];
    int64_t zeros = 0;
    for (int64_t chunk = s0; chunk < s1; chunk += 64) {
        const int64_t c = chunk + lane;
        const bool live = c < s1;
        const uint32_t code = live ? cell_code[c] : 0u;
        const int e0 = code ? lut_ptr[code] : 0;
        const int ne = code ? lut_ptr[code + 1] - e0 : 0;
        const uint32_t pending = ne >= 32 ? 0xffffffffu : ((1u << ne) - 1u);
        const int col = live ? (int)(cell_key[c] - panel_base) : 0;
        const int64_t f0 = live ? cell_first[c] : 0, f1 = live ? cell_first[c + 1] : 0;
        place_chunk(
            pending, [&](int j) { return lut_ids[e0 + j]; }, cur_c, cur_v, id_pattern,
            [&](int, int id, int pos_c, int pos_v) {
                cols[cbase + pos_c] = col;
                const unsigned pat = (unsigned)id_pattern[id];
                float *out = vals + vbase + pos_v;
                int kk = 0;
                int64_t f = f0;
                for (int r = 0; r < M_r; ++r) {
                    while (f < f1 && (int)ent_row[ent_idx[f]] < r) ++f;
                    if ((pat >> r) & 1u) {
                        const float v = (f < f1 && (int)ent_row[ent_idx[f]] == r) ? Lx[ent_idx[f]] : 0.0f;
                        zeros += (v == 0.0f);
                        out[kk++] = v;
                    }
                }
            });
    }
    zeros = wave_sum(zeros);
    if (lane == 0) padding[q] = zeros;
}

// One wave per CSR row: key = panel * k + column of every entry (entry e =
// position in Li minus Lp[0]), its row inside the panel; columns outside
// [0, k) set *bad.
template <typename I>
__global__ void cell_keys_kernel(int m, int64_t k, int M_r, const I *__restrict__ Lp, const I *__restrict__ Li,
                                 uint64_t *__restrict__ keys, uint32_t *__restrict__ idx,
                                 uint8_t *__restrict__ ent_row, int *__restrict__ bad) {
    const int row = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= m) return;
    const int p = row / M_r, r = row - p * M_r;
    const I base = Lp[0];
    const I beg = Lp[row], end = Lp[row + 1];
    for (I q = beg + lane; q < end; q += 64) {
        const I col = Li[q];
        const int64_t e = (int64_t)(q - base);
        if (col < 0 || (int64_t)col >= k) {
            atomicOr(bad, 1);
            keys[e] = 0;
        } else {
            keys[e] = (uint64_t)p * (uint64_t)k + (uint64_t)col;
        }
        idx[e] = (uint32_t)e;
        ent_row[e] = (uint8_t)r;
    }
}

// head[i] = 1 where a new cell starts in the sorted keys; two entries of one
// (row, column) set *bad = 2 (a cell's entries are adjacent, in row order).
__global__ void cell_heads_kernel(int64_t n, const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx,
                                  const uint8_t *__restrict__ ent_row, uint32_t *__restrict__ head,
                                  int *__restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool h = i == 0 || keys[i] != keys[i - 1];
    head[i] = h ? 1u : 0u;
    if (!h && ent_row[idx[i]] == ent_row[idx[i - 1]]) atomicOr(bad, 2);
}

// Per cell: key, first entry, code.  cell_first[ncells] = n (sentinel).
__global__ void cell_fill_kernel(int64_t n, const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx,
                                 const uint8_t *__restrict__ ent_row, const uint32_t *__restrict__ head,
                                 const uint32_t *__restrict__ cell_of, uint64_t *__restrict__ cell_key,
                                 uint32_t *__restrict__ cell_code, int64_t *__restrict__ cell_first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == n - 1) cell_first[cell_of[i] + head[i]] = n;
    if (!head[i]) return;
    const uint32_t c = cell_of[i];
    uint32_t code = 0;
    for (int64_t j = i; j < n && (j == i || !head[j]); ++j) code |= 1u << ent_row[idx[j]];
    cell_key[c] = keys[i];
    cell_code[c] = code;
    cell_first[c] = i;
}

// seg_cells[q] = first cell whose key >= the start of segment q = (p, t):
// p * k + t * Kc; seg_cells[Q] = ncells.
__global__ void segment_bounds_kernel(int Q, int KT, int64_t k, int Kc, int64_t ncells,
                                      const uint64_t *__restrict__ cell_key, int64_t *__restrict__ seg_cells) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > Q) return;
    if (q == Q) {
        seg_cells[q] = ncells;
        return;
    }
    const uint64_t want = (uint64_t)(q / KT) * (uint64_t)k + (uint64_t)(q % KT) * (uint64_t)Kc;
    int64_t lo = 0, hi = ncells;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (cell_key[mid] < want) lo = mid + 1; else hi = mid;
    }
    seg_cells[q] = lo;
}

// One thread per panel: number of ids with a nonzero count.
__global__ void group_count_kernel(int P, int NK, const int32_t *__restrict__ counts,
                                   int64_t *__restrict__ ngroups) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int32_t *c = counts + (int64_t)p * NK;
    int64_t g = 0;
    for (int i = 0; i < NK; ++i) g += (c[i] != 0);
    ngroups[p] = g;
}

// One thread per panel: (id | popcount(pattern) << 16, count) of every
// non-empty id, id order.
__global__ void group_fill_kernel(int P, int NK, const int32_t *__restrict__ counts,
                                  const int64_t *__restrict__ grp_ptr,
                                  const int32_t *__restrict__ id_pattern,
                                  int32_t *__restrict__ groups) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int32_t *c = counts + (int64_t)p * NK;
    int32_t *out = groups + 2 * grp_ptr[p];
    for (int i = 0; i < NK; ++i) {
        if (c[i]) {
            out[0] = i | (__popc((unsigned)id_pattern[i]) << 16);
            out[1] = c[i];
            out += 2;
        }
    }
}

// One thread per wave slot: where chunk j of kp of panel p starts and ends.
//   d0 = {p, group index, skip in group, first column}
//   d1 = {first value, end column, kp, j}      (panel-relative offsets)
// Chunk j covers packed columns [floor(j*n/kp), floor((j+1)*n/kp)).
__global__ void wave_desc_kernel(int nslots, const int32_t *__restrict__ slots,
                                 const int2 *__restrict__ groups,
                                 const int64_t *__restrict__ grp_ptr,
                                 const int64_t *__restrict__ col_ptr,
                                 const int32_t *__restrict__ id_pattern, int4 *__restrict__ desc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots) return;
    const int p = slots[3 * i], j = slots[3 * i + 1], kp = slots[3 * i + 2];
    if (p < 0) {
        desc[4 * i] = make_int4(-1, 0, 0, 0);
        desc[4 * i + 1] = make_int4(0, 0, 1, 0);
        return;
    }
    const int64_t ncols = col_ptr[p + 1] - col_ptr[p];
    const int64_t c0 = (int64_t)j * ncols / kp;
    const int64_t c1 = (int64_t)(j + 1) * ncols / kp;
    const int2 *gl = groups + grp_ptr[p];
    const int ng = (int)(grp_ptr[p + 1] - grp_ptr[p]);
    int64_t cs = 0, vs = 0;
    int g = 0;
    for (; g < ng; ++g) {
        const int2 e = gl[g];
        if (cs + e.y > c0) break;
        cs += e.y;
        vs += (int64_t)e.y * (e.x >> 16);
    }
    const int64_t skip = c0 - cs;
    const int pc = g < ng ? (gl[g].x >> 16) : 0;
    desc[4 * i] = make_int4(p, g, (int)skip, (int)c0);
    desc[4 * i + 1] = make_int4((int)(vs + skip * pc), (int)c1, kp, j);
}

// Batch records of every wave slot (one thread per slot).  Slot i's chunk is
// packed columns [c0, c1) of panel p (its descriptor); its records start at
// record index rec_start[i], the sum over the slots before it of their chunk
// lengths plus one, so no two runs overlap whatever order the slots hold the
// panels in (a run holds at most c1 - c0 batches and a terminator).  The
// start index replaces the descriptor's group skip (d0.z), which only this
// builder reads.
__global__ void wave_prog_kernel(int nslots, int U, int RS, const int2 *__restrict__ groups,
                                 const int64_t *__restrict__ grp_ptr, const int64_t *__restrict__ col_ptr,
                                 const int64_t *__restrict__ val_ptr, const int32_t *__restrict__ cols,
                                 const int64_t *__restrict__ rec_start, int4 *__restrict__ desc,
                                 int32_t *__restrict__ prog) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nslots) return;
    const int4 d0 = desc[4 * i], d1 = desc[4 * i + 1];
    const int p = d0.x;
    if (p < 0) return;
    const int64_t cbase = col_ptr[p];
    int64_t c = cbase + d0.w;
    const int64_t cend = cbase + d1.y;
    int64_t v = val_ptr[p] + d1.x;
    const int64_t start = rec_start[i];
    int32_t *rec = prog + start * RS;
    const int2 *g = groups + grp_ptr[p] + d0.y;
    int skip = d0.z;
    while (c < cend) {
        const int2 e = *g++;
        const int id = e.x & 0xffff, pc = e.x >> 16;
        int64_t take = e.y - skip;
        if (take > cend - c) take = cend - c;
        skip = 0;
        while (take > 0) {
            const int n = take < U ? (int)take : U;
            rec[0] = id * U + n - 1;
            rec[1] = (int)v;
            for (int u = 0; u < U; ++u) rec[2 + u] = cols[c + (u < n ? u : n - 1)];
            c += n;
            v += (int64_t)pc * n;
            take -= n;
            rec += RS;
        }
    }
    rec[0] = -1;  // terminator: its rows (column 0) and values (index v) stay valid reads
    rec[1] = (int)v;
    for (int u = 0; u < U; ++u) rec[2 + u] = 0;
    desc[4 * i] = make_int4(d0.x, d0.y, (int)start, d0.w);
    // The first record also sits in the descriptor's second half (when it
    // fits), so a wave issues its first B loads after one scalar load.
    if (2 + U <= 8) {
        int *w = reinterpret_cast<int *>(desc + 4 * i + 2);
        const int32_t *r0 = prog + start * RS;
        for (int k = 0; k < 2 + U; ++k) w[k] = r0[k];
    }
}

// ------------------------------------------------------- entry-list variants
__global__ void encode_entries_kernel(int n, int M_r, const int32_t *__restrict__ pattern_id,
                                      const int32_t *__restrict__ pat_codes,
                                      const int32_t *__restrict__ col_indices, int K,
                                      int32_t *__restrict__ ids, int32_t *__restrict__ bad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int pc = pat_codes[e];
    if (col_indices) {
        const int c = col_indices[e];
        if (c < 0 || c >= K) atomicOr(bad, 2);
    }
    int id;
    if (pc == 0) {
        id = SRT_ZERO_PATTERN_ID;
    } else if (pc < 0 || pc >= (1 << M_r) || pattern_id[pc] < 0) {
        atomicOr(bad, 1);
        id = SRT_ZERO_PATTERN_ID;
    } else {
        id = pattern_id[pc];
    }
    ids[e] = id;
}

// counts[NK + 1]: the last bucket counts zero patterns.
__global__ void count_entries_kernel(int n, int NK, const int32_t *__restrict__ ids,
                                     const int32_t *__restrict__ id_pattern,
                                     int32_t *__restrict__ counts, int64_t *__restrict__ ncols,
                                     int64_t *__restrict__ nvals) {
    extern __shared__ int hist[];
    __shared__ unsigned long long tot_c, tot_v;
    for (int i = threadIdx.x; i <= NK; i += blockDim.x) hist[i] = 0;
    if (threadIdx.x == 0) tot_c = tot_v = 0;
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const int id = ids[e];
        atomicAdd(&hist[id == SRT_ZERO_PATTERN_ID ? NK : id], 1);
    }
    __syncthreads();
    int64_t lc = 0, lv = 0;
    for (int i = threadIdx.x; i <= NK; i += blockDim.x) {
        const int h = hist[i];
        counts[i] = h;
        if (i < NK) {
            lc += h;
            lv += (int64_t)h * __popc((unsigned)id_pattern[i]);
        }
    }
    block_add_i64(&tot_c, lc);
    block_add_i64(&tot_v, lv);
    __syncthreads();
    if (threadIdx.x == 0) {
        *ncols = (int64_t)tot_c;
        *nvals = (int64_t)tot_v;
    }
}

__global__ void __launch_bounds__(64)
    fill_entries_kernel(int n, int col_offset, int K, int NK, int M_r,
                        const int32_t *__restrict__ id_pattern, const int32_t *__restrict__ ids,
                        const int32_t *__restrict__ col_indices,
                        const float *__restrict__ dense_panel, const int32_t *__restrict__ counts,
                        int32_t *__restrict__ cols, float *__restrict__ vals,
                        int64_t *__restrict__ padding) {
    extern __shared__ int cur[];
    int *cur_c = cur, *cur_v = cur + NK;
    const int lane = threadIdx.x & 63;
    id_cursors(NK, counts, id_pattern, cur_c, cur_v);
    __syncthreads();
    int64_t zeros = 0;
    for (int chunk = 0; chunk < n; chunk += 64) {
        const int e = chunk + lane;
        const int id = e < n ? ids[e] : SRT_ZERO_PATTERN_ID;
        const uint32_t pending = (id != SRT_ZERO_PATTERN_ID) ? 1u : 0u;
        const int col = e < n ? col_indices[e] : 0;
        place_chunk(
            pending, [&](int) { return id; }, cur_c, cur_v, id_pattern,
            [&](int, int lid, int pos_c, int pos_v) {
                cols[pos_c] = col + col_offset;
                const unsigned pat = (unsigned)id_pattern[lid];
                const float *d = dense_panel + (int64_t)col * M_r;
                int k = 0;
                for (int r = 0; r < M_r; ++r) {
                    if ((pat >> r) & 1u) {
                        const float v = d[r];
                        zeros += (v == 0.0f);
                        vals[pos_v + k++] = v;
                    }
                }
            });
    }
    zeros = wave_sum(zeros);
    if (lane == 0) *padding = zeros;
}

// Number of nonzero floats in x[0..n), single workgroup.
__global__ void count_nonzero_kernel(const float *__restrict__ x, int64_t n,
                                     int64_t *__restrict__ out) {
    __shared__ unsigned long long tot;
    if (threadIdx.x == 0) tot = 0;
    __syncthreads();
    int64_t c = 0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) c += (x[i] != 0.0f);
    atomicAdd(&tot, (unsigned long long)c);
    __syncthreads();
    if (threadIdx.x == 0) *out = (int64_t)tot;
}

// Stable order by id with the zero pattern last (sort_encoded_patterns,
// template:63-73): bucket NK holds the zero pattern.
__global__ void __launch_bounds__(64)
    stable_order_kernel(int n, int NK, const int32_t *__restrict__ ids,
                        const int32_t *__restrict__ counts, int32_t *__restrict__ encoded_sorted,
                        int32_t *__restrict__ order) {
    extern __shared__ int cur[];
    const int lane = threadIdx.x & 63;
    if (lane == 0) {
        int run = 0;
        for (int i = 0; i <= NK; ++i) {
            cur[i] = run;
            run += counts[i];
        }
    }
    __syncthreads();
    for (int chunk = 0; chunk < n; chunk += 64) {
        const int e = chunk + lane;
        const int id = e < n ? ids[e] : -1;
        const int bucket = id == SRT_ZERO_PATTERN_ID ? NK : id;
        uint32_t pending = e < n ? 1u : 0u;
        while (true) {
            const uint64_t active = __ballot(pending != 0);
            if (!active) break;
            const int leader = __ffsll((long long)active) - 1;
            const int lb = __shfl(bucket, leader, 64);
            const bool mine = pending && bucket == lb;
            const uint64_t m = __ballot(mine);
            const int rank = __popcll(m & lanemask_lt());
            const int base = cur[lb];
            if (mine) {
                pending = 0;
                encoded_sorted[base + rank] = id;
                order[base + rank] = e;
            }
            __syncthreads();
            if (lane == 0) cur[lb] = base + __popcll(m);
            __syncthreads();
        }
    }
}

}  // namespace

hipError_t launch_cell_keys(int m, int64_t k, int M_r, const void *Lp, const void *Li, int index_bytes,
                            uint64_t *keys, uint32_t *idx, uint8_t *ent_row, int *bad, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    const int waves = 4;
    const dim3 grid((m + waves - 1) / waves), block(64 * waves);
    if (index_bytes == 8)
        hipLaunchKernelGGL(cell_keys_kernel<int64_t>, grid, block, 0, s, m, k, M_r, (const int64_t *)Lp,
                           (const int64_t *)Li, keys, idx, ent_row, bad);
    else
        hipLaunchKernelGGL(cell_keys_kernel<int32_t>, grid, block, 0, s, m, k, M_r, (const int32_t *)Lp,
                           (const int32_t *)Li, keys, idx, ent_row, bad);
    return hipGetLastError();
}

hipError_t sort_cell_keys(int64_t n, int end_bit, uint64_t *keys, uint64_t *keys_alt, uint32_t *idx,
                          uint32_t *idx_alt, void *temp, size_t *temp_bytes, bool *sorted_in_alt, hipStream_t s) {
    hipcub::DoubleBuffer<uint64_t> kb(keys, keys_alt);
    hipcub::DoubleBuffer<uint32_t> vb(idx, idx_alt);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, kb, vb, (int)n, 0, end_bit, s);
    if (temp && e == hipSuccess) *sorted_in_alt = kb.Current() == keys_alt;
    return e;
}

hipError_t launch_cell_heads(int64_t n, const uint64_t *keys, const uint32_t *idx, const uint8_t *ent_row,
                             uint32_t *head, int *bad, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(cell_heads_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, keys, idx,
                       ent_row, head, bad);
    return hipGetLastError();
}

hipError_t scan_cell_heads(int64_t n, const uint32_t *head, uint32_t *cell_of, void *temp, size_t *temp_bytes,
                           hipStream_t s) {
    return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, head, cell_of, (int)n, s);
}

hipError_t launch_cell_fill(int64_t n, const uint64_t *keys, const uint32_t *idx, const uint8_t *ent_row,
                            const uint32_t *head, const uint32_t *cell_of, uint64_t *cell_key, uint32_t *cell_code,
                            int64_t *cell_first, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(cell_fill_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, keys, idx, ent_row,
                       head, cell_of, cell_key, cell_code, cell_first);
    return hipGetLastError();
}

hipError_t launch_segment_bounds(int Q, int KT, int64_t k, int Kc, int64_t ncells, const uint64_t *cell_key,
                                 int64_t *seg_cells, hipStream_t s) {
    hipLaunchKernelGGL(segment_bounds_kernel, dim3((Q + 256) / 256), dim3(256), 0, s, Q, KT, k, Kc, ncells,
                       cell_key, seg_cells);
    return hipGetLastError();
}

hipError_t launch_count_cells(int Q, const DeviceLut &lut, const int64_t *seg_cells, const uint32_t *cell_code,
                              int32_t *counts, int64_t *ncols, int64_t *nvals, hipStream_t s) {
    if (Q <= 0) return hipSuccess;
    hipLaunchKernelGGL(count_cells_kernel, dim3(Q), dim3(256), sizeof(int) * lut.num_patterns, s,
                       lut.num_patterns, lut.lut_ptr, lut.lut_ids, lut.id_pattern, seg_cells, cell_code,
                       counts, ncols, nvals);
    return hipGetLastError();
}

hipError_t launch_fill_cells(int Q, int64_t k, int KT, const DeviceLut &lut, const int64_t *seg_cells,
                             const uint64_t *cell_key, const uint32_t *cell_code, const int64_t *cell_first,
                             const uint32_t *ent_idx, const uint8_t *ent_row, const float *Lx,
                             const int64_t *col_ptr, const int64_t *val_ptr, const int32_t *counts,
                             int32_t *cols, float *vals, int64_t *padding, hipStream_t s) {
    if (Q <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_cells_kernel, dim3(Q), dim3(64), 2 * sizeof(int) * lut.num_patterns, s, k, KT,
                       lut.num_patterns, lut.M_r, lut.lut_ptr, lut.lut_ids, lut.id_pattern, seg_cells, cell_key,
                       cell_code, cell_first, ent_idx, ent_row, Lx, col_ptr, val_ptr, counts, cols, vals,
                       padding);
    return hipGetLastError();
}

hipError_t launch_exclusive_scan(int n, const int64_t *in, int64_t *out, hipStream_t s) {
    hipLaunchKernelGGL(exclusive_scan_kernel, dim3(1), dim3(1024), 0, s, n, in, out);
    return hipGetLastError();
}

hipError_t launch_group_count(int P, int NK, const int32_t *counts, int64_t *ngroups,
                              hipStream_t s) {
    if (P <= 0) return hipSuccess;
    hipLaunchKernelGGL(group_count_kernel, dim3((P + 255) / 256), dim3(256), 0, s, P, NK, counts,
                       ngroups);
    return hipGetLastError();
}

hipError_t launch_group_fill(int P, int NK, const int32_t *counts, const int64_t *grp_ptr,
                             const int32_t *id_pattern, int32_t *groups, hipStream_t s) {
    if (P <= 0) return hipSuccess;
    hipLaunchKernelGGL(group_fill_kernel, dim3((P + 255) / 256), dim3(256), 0, s, P, NK, counts,
                       grp_ptr, id_pattern, groups);
    return hipGetLastError();
}

hipError_t launch_wave_prog(int nslots, int U, const int32_t *groups, const int64_t *grp_ptr,
                            const int64_t *col_ptr, const int64_t *val_ptr, const int32_t *cols,
                            const int64_t *rec_start, int32_t *desc, int32_t *prog, hipStream_t s) {
    if (nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL(wave_prog_kernel, dim3((nslots + 255) / 256), dim3(256), 0, s, nslots, U, rec_stride(U),
                       reinterpret_cast<const int2 *>(groups), grp_ptr, col_ptr, val_ptr, cols, rec_start,
                       reinterpret_cast<int4 *>(desc), prog);
    return hipGetLastError();
}

hipError_t launch_wave_desc(int nslots, const int32_t *slots, const int32_t *groups,
                            const int64_t *grp_ptr, const int64_t *col_ptr,
                            const int32_t *id_pattern, int32_t *desc, hipStream_t s) {
    if (nslots <= 0) return hipSuccess;
    hipLaunchKernelGGL(wave_desc_kernel, dim3((nslots + 255) / 256), dim3(256), 0, s, nslots, slots,
                       reinterpret_cast<const int2 *>(groups), grp_ptr, col_ptr, id_pattern,
                       reinterpret_cast<int4 *>(desc));
    return hipGetLastError();
}

hipError_t launch_encode_entries(int n, const DeviceLut &lut, const int32_t *pat_codes,
                                 const int32_t *col_indices, int K, int32_t *ids, int32_t *bad,
                                 hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(encode_entries_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, lut.M_r,
                       lut.pattern_id, pat_codes, col_indices, K, ids, bad);
    return hipGetLastError();
}

hipError_t launch_count_nonzero(const float *x, int64_t n, int64_t *out, hipStream_t s) {
    hipLaunchKernelGGL(count_nonzero_kernel, dim3(1), dim3(1024), 0, s, x, n, out);
    return hipGetLastError();
}

hipError_t launch_count_entries(int n, const DeviceLut &lut, const int32_t *ids, int32_t *counts,
                                int64_t *ncols, int64_t *nvals, hipStream_t s) {
    hipLaunchKernelGGL(count_entries_kernel, dim3(1), dim3(256),
                       sizeof(int) * (lut.num_patterns + 1), s, n, lut.num_patterns, ids,
                       lut.id_pattern, counts, ncols, nvals);
    return hipGetLastError();
}

hipError_t launch_fill_entries(int n, int col_offset, int K, const DeviceLut &lut,
                               const int32_t *ids, const int32_t *col_indices,
                               const float *dense_panel, const int32_t *counts, int32_t *cols,
                               float *vals, int64_t *padding, hipStream_t s) {
    hipLaunchKernelGGL(fill_entries_kernel, dim3(1), dim3(64), 2 * sizeof(int) * lut.num_patterns,
                       s, n, col_offset, K, lut.num_patterns, lut.M_r, lut.id_pattern, ids,
                       col_indices, dense_panel, counts, cols, vals, padding);
    return hipGetLastError();
}

hipError_t launch_stable_order(int n, int NK, const int32_t *ids, const int32_t *counts,
                               int32_t *encoded_sorted, int32_t *order, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(stable_order_kernel, dim3(1), dim3(64), sizeof(int) * (NK + 1), s, n, NK,
                       ids, counts, encoded_sorted, order);
    return hipGetLastError();
}

}  // namespace srt
